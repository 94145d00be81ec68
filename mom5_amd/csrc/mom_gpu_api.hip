// C-ABI entry points of libmom5_gpu (declared in include/mom5_gpu.h):
// argument checking, context lifetime and dispatch to the kernels.
#include "mom_internal.h"

#include <cstdlib>
#include <cstring>

static thread_local char g_err[512] = "";

int mom_fail(mom_gpu_ctx *ctx, int code, const char *fmt, ...) {
  char *buf = ctx ? ctx->err : g_err;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, 512, fmt, ap);
  va_end(ap);
  if (ctx) std::snprintf(g_err, sizeof g_err, "%s", ctx->err);
  return code;
}

#define REQUIRE(ctx, cond, ...)                                     \
  do {                                                              \
    if (!(cond)) return mom_fail((ctx), MOM_GPU_ERR_ARG, __VA_ARGS__); \
  } while (0)

void side_join(mom_gpu_ctx *c) {
  if (c->side_pending) {
    (void)hipStreamWaitEvent(c->stream, c->side_done, 0);
    c->side_pending = false;
  }
  c->fork_valid = false;
}

extern "C" {

const char *mom_gpu_last_error(const mom_gpu_ctx_t *ctx) { return ctx ? ctx->err : g_err; }

int mom_gpu_malloc(void **ptr, size_t bytes) {
  if (!ptr) return mom_fail(nullptr, MOM_GPU_ERR_ARG, "mom_gpu_malloc: ptr is NULL");
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess)
    return mom_fail(nullptr, MOM_GPU_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes,
                    hipGetErrorString(e));
  return MOM_GPU_OK;
}

int mom_gpu_free(void *ptr) {
  hipError_t e = hipFree(ptr);
  if (e != hipSuccess) return mom_fail(nullptr, MOM_GPU_ERR_HIP, "hipFree: %s", hipGetErrorString(e));
  return MOM_GPU_OK;
}

// The copies wait for all work on the device first: contexts launch on
// their own non-blocking streams, which the null stream does not wait for.
int mom_gpu_memcpy_htod(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return mom_fail(nullptr, MOM_GPU_ERR_HIP, "hipMemcpy H2D: %s", hipGetErrorString(e));
  return MOM_GPU_OK;
}

int mom_gpu_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return mom_fail(nullptr, MOM_GPU_ERR_HIP, "hipMemcpy D2H: %s", hipGetErrorString(e));
  return MOM_GPU_OK;
}

int mom_gpu_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return mom_fail(nullptr, MOM_GPU_ERR_HIP, "hipMemcpy D2D: %s", hipGetErrorString(e));
  return MOM_GPU_OK;
}

int mom_gpu_memset(void *ptr, int value, size_t bytes) {
  // null-stream memset, completed before returning so that work later
  // enqueued on any context stream sees it
  hipError_t e = hipMemset(ptr, value, bytes);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return mom_fail(nullptr, MOM_GPU_ERR_HIP, "hipMemset: %s", hipGetErrorString(e));
  return MOM_GPU_OK;
}

static int *tuning_slot(mom_gpu_ctx *ctx, const char *knob, int *lo, int *hi) {
  if (!knob) return nullptr;
  if (!std::strcmp(knob, "barotropic_subcycles_per_launch")) {
    *lo = 0; *hi = 12;
    return &ctx->tune.bt_subcycles_per_launch;
  }
  if (!std::strcmp(knob, "mdfl_rows_per_wave")) {
    *lo = 1; *hi = 4;
    return &ctx->tune.mdfl_rows_per_wave;
  }
  if (!std::strcmp(knob, "step_group_f64")) {
    *lo = 1; *hi = 4;
    return &ctx->tune.step_group_f64;
  }
  if (!std::strcmp(knob, "step_group_f32")) {
    *lo = 1; *hi = 10;
    return &ctx->tune.step_group_f32;
  }
  if (!std::strcmp(knob, "step_rows_per_wave")) {
    *lo = 1; *hi = 4;
    return &ctx->tune.step_rows_per_wave;
  }
  if (!std::strcmp(knob, "step_rows_per_wave_f32")) {
    *lo = 1; *hi = 4;
    return &ctx->tune.step_rows_per_wave_f32;
  }
  if (!std::strcmp(knob, "step_waves_f32")) {
    *lo = 8; *hi = 16;
    return &ctx->tune.step_waves_f32;
  }
  if (!std::strcmp(knob, "step_waves_f64")) {
    *lo = 8; *hi = 16;
    return &ctx->tune.step_waves_f64;
  }
  if (!std::strcmp(knob, "vdiff_prefetch")) {
    *lo = 0; *hi = 1;
    return &ctx->tune.vdiff_prefetch;
  }
  if (!std::strcmp(knob, "vdiff_block")) {
    *lo = 64; *hi = 256;
    return &ctx->tune.vdiff_block;
  }
  if (!std::strcmp(knob, "vdiff_merge")) {
    *lo = 0; *hi = 1;
    return &ctx->tune.vdiff_merge;
  }
  if (!std::strcmp(knob, "vdiff_solver")) {
    *lo = 0; *hi = 1;
    return &ctx->tune.vdiff_solver;
  }
  if (!std::strcmp(knob, "step_streams")) {
    *lo = 1; *hi = 2;
    return &ctx->tune.step_streams;
  }
  if (!std::strcmp(knob, "barotropic_land_tiles")) {
    *lo = 0; *hi = 1;
    return &ctx->tune.bt_land_tiles;
  }
  return nullptr;
}

int mom_gpu_set_tuning(mom_gpu_ctx_t *ctx, const char *knob, int value) {
  REQUIRE(nullptr, ctx, "mom_gpu_set_tuning: NULL context");
  int lo = 0, hi = 0;
  int *slot = tuning_slot(ctx, knob, &lo, &hi);
  REQUIRE(ctx, slot, "mom_gpu_set_tuning: unknown knob '%s'", knob ? knob : "(null)");
  REQUIRE(ctx, value >= lo && value <= hi, "mom_gpu_set_tuning: %s=%d outside [%d,%d]", knob,
          value, lo, hi);
  REQUIRE(ctx, slot != &ctx->tune.mdfl_rows_per_wave || value == 1 || value == 2 || value == 4,
          "mom_gpu_set_tuning: %s must be 1, 2 or 4", knob);
  REQUIRE(ctx, (slot != &ctx->tune.step_rows_per_wave_f32 && slot != &ctx->tune.step_rows_per_wave) ||
                   value == 1 || value == 2 || value == 4,
          "mom_gpu_set_tuning: %s must be 1, 2 or 4", knob);
  REQUIRE(ctx, slot != &ctx->tune.vdiff_block || value == 64 || value == 256,
          "mom_gpu_set_tuning: vdiff_block must be 64 or 256");
  REQUIRE(ctx, slot != &ctx->tune.step_group_f64 || value != 3,
          "mom_gpu_set_tuning: step_group_f64 must be 1, 2 or 4");
  REQUIRE(ctx, slot != &ctx->tune.step_waves_f32 || value == 8 || value == 12 || value == 16,
          "mom_gpu_set_tuning: step_waves_f32 must be 8, 12 or 16");
  REQUIRE(ctx, slot != &ctx->tune.step_waves_f64 || value == 8 || value == 12 || value == 16,
          "mom_gpu_set_tuning: step_waves_f64 must be 8, 12 or 16");
  REQUIRE(ctx, slot != &ctx->tune.step_group_f32 || value == 1 || value == 2 || value == 4 ||
                   value == 5 || value == 6 || value == 8 || value == 10,
          "mom_gpu_set_tuning: step_group_f32 must be 1, 2, 4, 5, 6, 8 or 10");
  *slot = value;
  return MOM_GPU_OK;
}

int mom_gpu_get_tuning(const mom_gpu_ctx_t *ctx, const char *knob, int *value) {
  REQUIRE(nullptr, ctx && value, "mom_gpu_get_tuning: NULL argument");
  int lo = 0, hi = 0;
  int *slot = tuning_slot(const_cast<mom_gpu_ctx *>(ctx), knob, &lo, &hi);
  if (!slot) return mom_fail(const_cast<mom_gpu_ctx *>(ctx), MOM_GPU_ERR_ARG,
                             "mom_gpu_get_tuning: unknown knob '%s'", knob ? knob : "(null)");
  *value = *slot;
  return MOM_GPU_OK;
}

int mom_gpu_fill(mom_gpu_ctx_t *ctx, double *ptr, double value, size_t n) {
  REQUIRE(nullptr, ctx, "mom_gpu_fill: NULL context");
  side_join(ctx);
  REQUIRE(ctx, ptr || n == 0, "mom_gpu_fill: NULL pointer");
  return launch_fill(ctx, ptr, value, n);
}

int mom_gpu_update_domains(mom_gpu_ctx_t *ctx, double *field, int nz) {
  REQUIRE(nullptr, ctx, "mom_gpu_update_domains: NULL context");
  side_join(ctx);
  REQUIRE(ctx, field || nz == 0, "mom_gpu_update_domains: NULL field");
  REQUIRE(ctx, nz >= 0 && nz <= 65535, "mom_gpu_update_domains: nz=%d outside [0, 65535]", nz);
  if (nz == 0) return MOM_GPU_OK;
  return launch_update_domains(ctx, field, nz);
}

int mom_gpu_advection_velocity(mom_gpu_ctx_t *ctx, const double *u, const double *rho_dzu,
                               const double *pme, const double *river,
                               const double *rho_dzt_tendency, const double *mass_source,
                               double *uhrho_et, double *vhrho_nt, double *wrho_bt,
                               double *diverge_t) {
  REQUIRE(nullptr, ctx, "mom_gpu_advection_velocity: NULL context");
  side_join(ctx);
  REQUIRE(ctx, u && rho_dzu && pme && river && uhrho_et && vhrho_nt && wrho_bt,
          "mom_gpu_advection_velocity: NULL required array");
  return launch_adv_vel(ctx, u, rho_dzu, pme, river, rho_dzt_tendency, mass_source, uhrho_et,
                        vhrho_nt, wrho_bt, diverge_t);
}

int mom_gpu_barotropic_forcing(mom_gpu_ctx_t *ctx, const double *accel, const double *smf,
                               const double *bmf, double *forcing_bt) {
  REQUIRE(nullptr, ctx, "mom_gpu_barotropic_forcing: NULL context");
  REQUIRE(ctx, accel && smf && bmf && forcing_bt, "mom_gpu_barotropic_forcing: NULL argument");
  return launch_bt_forcing(ctx, accel, smf, bmf, forcing_bt);
}

int mom_gpu_eta_and_pbot_update(mom_gpu_ctx_t *ctx, int pressure_based, double rho0,
                                double grav, double dtime, const double *conv_rho_ud_t_tau,
                                const double *pme, const double *river, const double *source,
                                const double *dpatm_dt, const double *ht, const double *x_taum1,
                                double *x_taup1, double *pbot_t_taup1, double *tendency) {
  REQUIRE(nullptr, ctx, "mom_gpu_eta_and_pbot_update: NULL context");
  REQUIRE(ctx, conv_rho_ud_t_tau && pme && river && source && x_taum1 && x_taup1 && tendency,
          "mom_gpu_eta_and_pbot_update: NULL argument");
  REQUIRE(ctx, !pressure_based || (dpatm_dt && ht && pbot_t_taup1),
          "mom_gpu_eta_and_pbot_update: the pressure-based step needs dpatm_dt, ht and pbot_t");
  REQUIRE(ctx, (const double *)x_taup1 != x_taum1 && (const double *)tendency != x_taum1,
          "mom_gpu_eta_and_pbot_update: x_taup1 and tendency must not alias x_taum1");
  REQUIRE(ctx, rho0 > 0.0, "mom_gpu_eta_and_pbot_update: rho0 must be positive");
  return launch_eta_update(ctx, pressure_based ? 1 : 0, rho0, grav, dtime, conv_rho_ud_t_tau, pme,
                           river, source, dpatm_dt, ht, x_taum1, x_taup1, pbot_t_taup1, tendency);
}

int mom_gpu_set_remap_metrics(mom_gpu_ctx_t *ctx, const double *dte, const double *dtw,
                              const double *dus, const double *dun, const double *daur) {
  REQUIRE(nullptr, ctx, "mom_gpu_set_remap_metrics: NULL context");
  REQUIRE(ctx, dte && dtw && dus && dun && daur, "mom_gpu_set_remap_metrics: NULL argument");
  const size_t P2 = (size_t)ctx->kd.P2;
  if (!ctx->remap) MOM_HIP(ctx, hipMalloc(&ctx->remap, 5 * P2 * sizeof(double)));
  const double *src[5] = {dte, dtw, dus, dun, daur};
  for (int n = 0; n < 5; ++n)
    MOM_HIP(ctx, hipMemcpyAsync(ctx->remap + n * P2, src[n], P2 * sizeof(double),
                                hipMemcpyDefault, ctx->stream));
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MOM_GPU_OK;
}

int mom_gpu_barotropic_closure(mom_gpu_ctx_t *ctx, int pressure_based, double rho0, double grav,
                               const double *eta_t_bar, const double *anompb_bar,
                               const double *patm, const double *rho_g, const double *ht,
                               const double *tmasken, double *eta_u, double *ps,
                               double *grad_ps, double *grad_anompb) {
  REQUIRE(nullptr, ctx, "mom_gpu_barotropic_closure: NULL context");
  REQUIRE(ctx, ctx->remap,
          "mom_gpu_barotropic_closure: REMAP_BT_TO_BU metrics not set (mom_gpu_set_remap_metrics)");
  REQUIRE(ctx, eta_t_bar && patm && rho_g && eta_u && ps && grad_ps,
          "mom_gpu_barotropic_closure: NULL argument");
  REQUIRE(ctx, !pressure_based || (ht && anompb_bar),
          "mom_gpu_barotropic_closure: the pressure-based closure needs ht and anompb_bar");
  REQUIRE(ctx, !grad_anompb || anompb_bar,
          "mom_gpu_barotropic_closure: grad_anompb needs anompb_bar");
  return launch_bt_closure(ctx, pressure_based ? 1 : 0, rho0, grav, eta_t_bar, anompb_bar, patm,
                           rho_g, ht, tmasken, eta_u, ps, grad_ps, grad_anompb);
}

int mom_gpu_update_ocean_barotropic_variant(mom_gpu_ctx_t *ctx, const mom_gpu_bt_params_t *p,
                                            const mom_gpu_bt_variant_t *v) {
  REQUIRE(nullptr, ctx, "mom_gpu_update_ocean_barotropic_variant: NULL context");
  REQUIRE(ctx, p && v, "mom_gpu_update_ocean_barotropic_variant: NULL argument");
  REQUIRE(ctx, p->nts >= 1, "mom_gpu_update_ocean_barotropic_variant: nts=%d < 1", p->nts);
  REQUIRE(ctx, p->rho0 > 0.0, "mom_gpu_update_ocean_barotropic_variant: rho0 must be positive");
  REQUIRE(ctx, v->x_bar_tau && v->udrho_tau && v->forcing_bt && v->thick && v->pme &&
                   v->river && v->source && v->smooth && v->x_bar_taup1 && v->udrho_taup1,
          "mom_gpu_update_ocean_barotropic_variant: NULL field");
  REQUIRE(ctx, p->initsum_with_bar || v->x_tau,
          "mom_gpu_update_ocean_barotropic_variant: x_tau needed when initsum_with_bar = 0");
  REQUIRE(ctx, !v->cgrid || v->tmasken,
          "mom_gpu_update_ocean_barotropic_variant: the C-grid needs tmasken");
  REQUIRE(ctx, v->pressure_based ? v->dpatm_dt != nullptr : (v->patm && v->rho_g),
          "mom_gpu_update_ocean_barotropic_variant: %s",
          v->pressure_based ? "the pressure-based scheme needs dpatm_dt"
                            : "the depth-based scheme needs patm and rho_g");
  const void *outs[] = {v->x_bar_taup1, v->udrho_taup1, v->conv_rho_ud_t};
  const void *ins[] = {v->x_tau, v->x_bar_tau, v->udrho_tau};
  for (const void *o : outs)
    for (const void *i : ins)
      REQUIRE(ctx, !o || o != i,
              "mom_gpu_update_ocean_barotropic_variant: an output aliases x_tau, x_bar_tau or udrho_tau");
  return launch_bt_variant(ctx, p, v);
}

int mom_gpu_copy(mom_gpu_ctx_t *ctx, double *dst, const double *src, size_t n) {
  REQUIRE(nullptr, ctx, "mom_gpu_copy: NULL context");
  side_join(ctx);
  REQUIRE(ctx, (dst && src) || n == 0, "mom_gpu_copy: NULL pointer");
  return launch_copy(ctx, dst, src, n);
}

static void free_ctx(mom_gpu_ctx *c) {
  if (!c) return;
  double **ds[] = {&c->dxt, &c->dyt, &c->dat, &c->datr, &c->dxte, &c->dyte, &c->dxtn, &c->dytn,
                   &c->dxu, &c->dyu, &c->dxur, &c->dyur, &c->f, &c->dzt, &c->qw, &c->qzw,
                   &c->bt, &c->tri_e, &c->tri_e2, &c->red, &c->remap, &c->btv};
  for (double **p : ds)
    if (*p) hipFree(*p);
  if (c->kmt) hipFree(c->kmt);
  if (c->kmu) hipFree(c->kmu);
  if (c->vd_scr) hipFree(c->vd_scr);
  for (int i = 0; i < 8; ++i)
    if (c->ev[i]) hipEventDestroy(c->ev[i]);
  if (c->side_done) hipEventDestroy(c->side_done);
  if (c->fork_ev) hipEventDestroy(c->fork_ev);
  if (c->join_ev) hipEventDestroy(c->join_ev);
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int mom_gpu_create(const mom_gpu_domain_t *dom, const mom_gpu_grid_t *grid, mom_gpu_ctx_t **out) {
  if (!dom || !grid || !out) return mom_fail(nullptr, MOM_GPU_ERR_ARG, "mom_gpu_create: NULL argument");
  *out = nullptr;
  const int ni = dom->iec - dom->isc + 1, nj = dom->jec - dom->jsc + 1;
  if (ni < 1 || nj < 1 || dom->nk < 2)
    return mom_fail(nullptr, MOM_GPU_ERR_ARG, "mom_gpu_create: empty domain (ni=%d nj=%d nk=%d; nk>=2 required)",
                    ni, nj, dom->nk);
  if (dom->isd != dom->isc - 1 || dom->ied != dom->iec + 1 || dom->jsd != dom->jsc - 1 ||
      dom->jed != dom->jec + 1)
    return mom_fail(nullptr, MOM_GPU_ERR_ARG,
                    "mom_gpu_create: only a one-cell halo is supported (isd=isc-1, ied=iec+1, ...)");
  const long long cells = (long long)(ni + 2) * (nj + 2) * (dom->nk + 1);
  if (cells >= (1LL << 31))
    return mom_fail(nullptr, MOM_GPU_ERR_ARG, "mom_gpu_create: domain too large for 32-bit offsets");
  const double *need[] = {grid->dxt, grid->dyt, grid->dat, grid->datr, grid->dxte, grid->dyte,
                          grid->dxtn, grid->dytn, grid->dxu, grid->dyu, grid->dxur, grid->dyur,
                          grid->f, grid->dzt};
  for (const double *p : need)
    if (!p) return mom_fail(nullptr, MOM_GPU_ERR_ARG, "mom_gpu_create: a grid field is NULL");
  if (!grid->kmt || !grid->kmu) return mom_fail(nullptr, MOM_GPU_ERR_ARG, "mom_gpu_create: kmt/kmu NULL");

  mom_gpu_ctx *c = new mom_gpu_ctx();
  std::memset(c, 0, sizeof *c);
  c->dom = *dom;
  KDom &k = c->kd;
  k.ni = ni; k.nj = nj; k.nk = dom->nk;
  k.NI = ni + 2; k.NJ = nj + 2;
  k.P2 = k.NI * k.NJ;
  k.cyclic_x = dom->cyclic_x ? 1 : 0;
  k.cyclic_y = dom->cyclic_y ? 1 : 0;
  c->tune.bt_subcycles_per_launch = 5;
  c->tune.mdfl_rows_per_wave = 1;
  c->tune.step_group_f64 = 2;
  c->tune.step_group_f32 = 10;
  c->tune.step_rows_per_wave = 1;
  c->tune.step_rows_per_wave_f32 = 1;
  c->tune.step_waves_f32 = 16;
  c->tune.step_waves_f64 = 16;
  c->tune.vdiff_prefetch = 1;
  c->tune.vdiff_block = 256;
  c->tune.vdiff_merge = 0;
  c->tune.vdiff_solver = 1;
  c->tune.step_streams = 1;
  c->tune.bt_land_tiles = 1;
  const size_t P2 = (size_t)k.P2, nk = (size_t)k.nk;

#define CK(call)                                                                      \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      int rc_ = mom_fail(nullptr, MOM_GPU_ERR_HIP, "mom_gpu_create: %s: %s", #call,   \
                         hipGetErrorString(e_));                                      \
      free_ctx(c);                                                                    \
      return rc_;                                                                     \
    }                                                                                 \
  } while (0)
  CK(hipGetDevice(&c->device));
  CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int i = 0; i < 8; ++i) CK(hipEventCreate(&c->ev[i]));
  CK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  CK(hipEventCreateWithFlags(&c->side_done, hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
  CK(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
  struct {
    double **dst;
    const double *src;
  } f2[] = {{&c->dxt, grid->dxt}, {&c->dyt, grid->dyt}, {&c->dat, grid->dat}, {&c->datr, grid->datr},
            {&c->dxte, grid->dxte}, {&c->dyte, grid->dyte}, {&c->dxtn, grid->dxtn}, {&c->dytn, grid->dytn},
            {&c->dxu, grid->dxu}, {&c->dyu, grid->dyu}, {&c->dxur, grid->dxur}, {&c->dyur, grid->dyur},
            {&c->f, grid->f}};
  for (auto &e : f2) {
    CK(hipMalloc(e.dst, P2 * sizeof(double)));
    CK(hipMemcpyAsync(*e.dst, e.src, P2 * sizeof(double), hipMemcpyDefault, c->stream));
  }
  CK(hipMalloc(&c->kmt, P2 * sizeof(int)));
  CK(hipMemcpyAsync(c->kmt, grid->kmt, P2 * sizeof(int), hipMemcpyDefault, c->stream));
  CK(hipMalloc(&c->kmu, P2 * sizeof(int)));
  CK(hipMemcpyAsync(c->kmu, grid->kmu, P2 * sizeof(int), hipMemcpyDefault, c->stream));
  CK(hipMalloc(&c->dzt, nk * sizeof(double)));
  CK(hipMemcpyAsync(c->dzt, grid->dzt, nk * sizeof(double), hipMemcpyDefault, c->stream));
  CK(hipMalloc(&c->qw, 16 * P2 * sizeof(double)));
  CK(hipMalloc(&c->qzw, 8 * nk * sizeof(double)));
  CK(hipMalloc(&c->bt, 9 * P2 * sizeof(double)));
  CK(hipMalloc(&c->tri_e, P2 * nk * sizeof(double)));
  c->red_n = 1024;
  CK(hipMalloc(&c->red, (8 * (size_t)c->red_n + 8) * sizeof(double)));
#undef CK

  long long bad = 0;
  int rc = MOM_GPU_OK;
  if (grid->tmask && grid->umask) rc = launch_check_masks(c, grid->tmask, grid->umask, &bad);
  if (rc == MOM_GPU_OK && bad != 0)
    rc = mom_fail(c, MOM_GPU_ERR_GRID,
                  "mom_gpu_create: %lld mask entries disagree with kmt/kmu "
                  "(expected tmask = k<=kmt, umask = k<=kmu, ocean_grids.F90:1374-1386)", bad);
  if (rc == MOM_GPU_OK) rc = launch_quicker_init(c);
  if (rc == MOM_GPU_OK) rc = launch_check_recip(c);
  if (rc == MOM_GPU_OK) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = mom_fail(c, MOM_GPU_ERR_HIP, "mom_gpu_create: %s", hipGetErrorString(e));
  }
  if (rc != MOM_GPU_OK) {
    std::snprintf(g_err, sizeof g_err, "%s", c->err);
    free_ctx(c);
    return rc;
  }
  *out = c;
  return MOM_GPU_OK;
}

int mom_gpu_destroy(mom_gpu_ctx_t *ctx) {
  if (!ctx) return MOM_GPU_OK;
  (void)hipStreamSynchronize(ctx->side);
  (void)hipStreamSynchronize(ctx->stream);
  free_ctx(ctx);
  return MOM_GPU_OK;
}

int mom_gpu_sync(mom_gpu_ctx_t *ctx) {
  REQUIRE(nullptr, ctx, "mom_gpu_sync: NULL context");
  side_join(ctx);
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MOM_GPU_OK;
}

int mom_gpu_mark(mom_gpu_ctx_t *ctx, int slot) {
  REQUIRE(nullptr, ctx, "mom_gpu_mark: NULL context");
  REQUIRE(ctx, slot >= 0 && slot < 8, "mom_gpu_mark: slot %d out of range", slot);
  side_join(ctx);
  MOM_HIP(ctx, hipEventRecord(ctx->ev[slot], ctx->stream));
  return MOM_GPU_OK;
}

int mom_gpu_elapsed(mom_gpu_ctx_t *ctx, int s0, int s1, float *ms) {
  REQUIRE(nullptr, ctx, "mom_gpu_elapsed: NULL context");
  REQUIRE(ctx, s0 >= 0 && s0 < 8 && s1 >= 0 && s1 < 8 && ms, "mom_gpu_elapsed: bad slots");
  MOM_HIP(ctx, hipEventSynchronize(ctx->ev[s1]));
  MOM_HIP(ctx, hipEventElapsedTime(ms, ctx->ev[s0], ctx->ev[s1]));
  return MOM_GPU_OK;
}

// ---- ensembles -----------------------------------------------------------------
}  // extern "C"

const void *ens_table(mom_gpu_ensemble *e, const void *host, size_t bytes) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(e->lead->stream, &cs) != hipSuccess) {
    mom_fail(e->lead, MOM_GPU_ERR_HIP, "ensemble: hipStreamIsCapturing failed");
    return nullptr;
  }
  const bool capturing = cs != hipStreamCaptureStatusNone;
  const int hit = e->tables.lookup(host, bytes, capturing);
  if (hit >= 0) return e->tables.entries[hit].dev;
  if (capturing) {
    mom_fail(e->lead, MOM_GPU_ERR_ARG,
             "ensemble: a new launch table while capturing; run one step of each time-level "
             "parity before capturing");
    return nullptr;
  }
  const int v = e->tables.victim();       // keep the cache bounded
  if (v == -2) {
    mom_fail(e->lead, MOM_GPU_ERR_ARG,
             "ensemble: %zu launch tables are held by captured graphs; destroy the ensemble "
             "(and its graphs) to release them", e->tables.entries.size());
    return nullptr;
  }
  if (v >= 0) {
    if (hipStreamSynchronize(e->lead->stream) != hipSuccess) return nullptr;
    (void)hipFree(e->tables.entries[v].dev);
    e->tables.erase(v);
  }
  void *dev = nullptr;
  hipError_t err = hipMalloc(&dev, bytes);
  if (err == hipSuccess) err = hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    if (dev) (void)hipFree(dev);
    mom_fail(e->lead, MOM_GPU_ERR_HIP, "ensemble table: %s", hipGetErrorString(err));
    return nullptr;
  }
  e->tables.add(host, bytes, dev);
  return dev;
}

// Ensemble calls run on the lead stream.  Outside a capture, the lead stream
// first waits for work already queued on every member's stream, and each
// member stream afterwards waits for the ensemble's work, so calls on a
// member context before and after an ensemble call stay ordered with it.
// During a capture the members must be idle (synchronised) instead: their
// streams take no part in the graph.
static int ens_capturing(mom_gpu_ensemble *e, bool *capturing) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  MOM_HIP(e->lead, hipStreamIsCapturing(e->lead->stream, &cs));
  *capturing = cs != hipStreamCaptureStatusNone;
  return MOM_GPU_OK;
}

static int ens_join_members(mom_gpu_ensemble *e, bool capturing) {
  for (mom_gpu_ctx *c : e->ctx) {
    side_join(c);                           // members' pending second-stream work
    if (capturing || c->stream == e->lead->stream) continue;
    MOM_HIP(e->lead, hipEventRecord(c->join_ev, c->stream));
    MOM_HIP(e->lead, hipStreamWaitEvent(e->lead->stream, c->join_ev, 0));
  }
  return MOM_GPU_OK;
}

static int ens_release_members(mom_gpu_ensemble *e, bool capturing) {
  if (capturing) return MOM_GPU_OK;
  MOM_HIP(e->lead, hipEventRecord(e->lead->join_ev, e->lead->stream));
  for (mom_gpu_ctx *c : e->ctx)
    if (c->stream != e->lead->stream)
      MOM_HIP(e->lead, hipStreamWaitEvent(c->stream, e->lead->join_ev, 0));
  return MOM_GPU_OK;
}

extern "C" {

int mom_gpu_ensemble_create(int ndom, mom_gpu_ctx_t *const *ctxs, mom_gpu_ensemble_t **out) {
  REQUIRE(nullptr, out && ctxs && ndom >= 1, "mom_gpu_ensemble_create: bad argument");
  *out = nullptr;
  REQUIRE(nullptr, ndom <= 64, "mom_gpu_ensemble_create: at most 64 domains (%d)", ndom);
  for (int d = 0; d < ndom; ++d) {
    REQUIRE(nullptr, ctxs[d], "mom_gpu_ensemble_create: NULL context %d", d);
    REQUIRE(ctxs[0], ctxs[d]->device == ctxs[0]->device,
            "mom_gpu_ensemble_create: contexts on different devices");
  }
  auto *e = new mom_gpu_ensemble();
  e->ctx.assign(ctxs, ctxs + ndom);
  e->lead = ctxs[0];
  if (hipMalloc((void **)&e->counters, sizeof(int) * mom_gpu_ensemble::NCOUNT) != hipSuccess) {
    delete e;
    return mom_fail(ctxs[0], MOM_GPU_ERR_NOMEM, "mom_gpu_ensemble_create: counters");
  }
  *out = e;
  return MOM_GPU_OK;
}

int mom_gpu_ensemble_destroy(mom_gpu_ensemble_t *e) {
  if (!e) return MOM_GPU_OK;
  (void)hipStreamSynchronize(e->lead->stream);
  for (auto &t : e->tables.entries) (void)hipFree(t.dev);
  if (e->counters) (void)hipFree(e->counters);
  delete e;
  return MOM_GPU_OK;
}

static int check_step(mom_gpu_ctx *ctx, const mom_gpu_tracer_step_t *s, const char *who);

int mom_gpu_ensemble_step_tracers(mom_gpu_ensemble_t *e, const mom_gpu_tracer_step_t *steps) {
  REQUIRE(nullptr, e && steps, "mom_gpu_ensemble_step_tracers: NULL argument");
  for (size_t d = 0; d < e->ctx.size(); ++d) {
    const int rc = check_step(e->ctx[d], &steps[d], "mom_gpu_ensemble_step_tracers");
    if (rc) return mom_fail(e->lead, rc, "domain %zu: %s", d, e->ctx[d]->err);
    if (steps[d].scheme)
      for (int n = 0; n < steps[d].ntr; ++n)
        REQUIRE(e->lead, steps[d].scheme[n] == MOM_ADVECT_MDFL_SWEBY,
                "mom_gpu_ensemble_step_tracers: domain %zu: ensembles take MDFL-Sweby tracers "
                "only (tracer %d)", d, n);
  }
  bool capturing = false;
  int rc = ens_capturing(e, &capturing);
  if (!rc) rc = ens_join_members(e, capturing);
  if (!rc) rc = launch_step_ensemble(e, steps);
  if (!rc) rc = launch_vdiff_ensemble(e, steps);
  if (!rc) rc = ens_release_members(e, capturing);
  return rc;
}

int mom_gpu_ensemble_update_barotropic(mom_gpu_ensemble_t *e, const mom_gpu_bt_params_t *p,
                                       const mom_gpu_bt_fields_t *f) {
  REQUIRE(nullptr, e && p && f, "mom_gpu_ensemble_update_barotropic: NULL argument");
  for (size_t d = 0; d < e->ctx.size(); ++d) {
    const mom_gpu_bt_fields_t &x = f[d];
    REQUIRE(e->lead, p[d].nts >= 1 && p[d].dtbt > 0.0 && p[d].rho0 > 0.0,
            "mom_gpu_ensemble_update_barotropic: domain %zu: bad parameters", d);
    REQUIRE(e->lead, e->ctx[d]->kd.cyclic_y == 0,
            "mom_gpu_ensemble_update_barotropic: domain %zu: cyclic_y unsupported", d);
    REQUIRE(e->lead, x.eta_t_bar_tau && x.udrho_tau && x.forcing_bt && x.thicku && x.patm &&
                         x.pme && x.river && x.source && x.eta_smooth && x.rho_g &&
                         x.eta_t_bar_taup1 && x.udrho_taup1 &&
                         (p[d].initsum_with_bar || x.eta_t_tau),
            "mom_gpu_ensemble_update_barotropic: domain %zu: NULL required field", d);
    const void *outs[] = {x.eta_t_bar_taup1, x.udrho_taup1, x.conv_rho_ud_t};
    const void *ins[] = {x.eta_t_tau, x.eta_t_bar_tau, x.udrho_tau};
    for (const void *o : outs)
      for (const void *i : ins)
        REQUIRE(e->lead, !o || o != i,
                "mom_gpu_ensemble_update_barotropic: domain %zu: an output aliases eta_t_tau, "
                "eta_t_bar_tau or udrho_tau", d);
  }
  bool capturing = false;
  int rc = ens_capturing(e, &capturing);
  if (!rc) rc = ens_join_members(e, capturing);
  if (!rc) rc = launch_barotropic_ensemble(e, p, f);
  if (!rc) rc = ens_release_members(e, capturing);
  return rc;
}

// ---- hipGraph capture of whole steps -----------------------------------------
struct mom_gpu_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
  int nodes;
};

int mom_gpu_graph_begin(mom_gpu_ctx_t *ctx) {
  REQUIRE(nullptr, ctx, "mom_gpu_graph_begin: NULL context");
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  MOM_HIP(ctx, hipStreamIsCapturing(ctx->stream, &st));
  REQUIRE(ctx, st == hipStreamCaptureStatusNone, "mom_gpu_graph_begin: already capturing");
  side_join(ctx);
  MOM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  return MOM_GPU_OK;
}

int mom_gpu_graph_end(mom_gpu_ctx_t *ctx, mom_gpu_graph_t **out) {
  REQUIRE(nullptr, ctx, "mom_gpu_graph_end: NULL context");
  REQUIRE(ctx, out, "mom_gpu_graph_end: NULL output");
  *out = nullptr;
  hipGraph_t g = nullptr;
  side_join(ctx);                 // the second stream rejoins the capture
  hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->fork_valid = false;        // events recorded in the capture are graph nodes
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    // streams whose capture was invalidated may stay unusable: replace them
    hipStream_t fresh = nullptr;
    if (hipStreamCreateWithFlags(&fresh, hipStreamNonBlocking) == hipSuccess) {
      (void)hipStreamDestroy(ctx->stream);
      ctx->stream = fresh;
    }
    fresh = nullptr;
    if (hipStreamCreateWithFlags(&fresh, hipStreamNonBlocking) == hipSuccess) {
      (void)hipStreamDestroy(ctx->side);
      ctx->side = fresh;
    }
    (void)hipGetLastError();
    return mom_fail(ctx, MOM_GPU_ERR_HIP,
                    "mom_gpu_graph_end: capture failed (%s); between begin and end only "
                    "kernel-launching entry points may be called (no sync, copies, marks, "
                    "content or first-use allocations)", hipGetErrorString(e));
  }
  size_t n = 0;
  (void)hipGraphGetNodes(g, nullptr, &n);
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return mom_fail(ctx, MOM_GPU_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  *out = new mom_gpu_graph{g, x, (int)n};
  return MOM_GPU_OK;
}

int mom_gpu_graph_launch(mom_gpu_ctx_t *ctx, const mom_gpu_graph_t *graph) {
  REQUIRE(nullptr, ctx, "mom_gpu_graph_launch: NULL context");
  REQUIRE(ctx, graph, "mom_gpu_graph_launch: NULL graph");
  side_join(ctx);
  MOM_HIP(ctx, hipGraphLaunch(graph->exec, ctx->stream));
  return MOM_GPU_OK;
}

int mom_gpu_graph_nodes(const mom_gpu_graph_t *graph) { return graph ? graph->nodes : -1; }

int mom_gpu_graph_destroy(mom_gpu_graph_t *graph) {
  if (!graph) return MOM_GPU_OK;
  (void)hipGraphExecDestroy(graph->exec);
  (void)hipGraphDestroy(graph->graph);
  delete graph;
  return MOM_GPU_OK;
}

static size_t p3bytes(const mom_gpu_ctx *c) { return (size_t)c->kd.P2 * c->kd.nk * sizeof(double); }

int mom_gpu_horz_advect_tracer(mom_gpu_ctx_t *ctx, int scheme, const double *T_taum1,
                               const double *T_tau, const double *uhrho_et,
                               const double *vhrho_nt, const double *wrho_bt,
                               const double *rho_dzt_tau, const double *tmask_limit,
                               int limit_with_upwind, double dtime, double *th_tendency,
                               double *wrk1, double *flux_x, double *flux_y, double *flux_z) {
  REQUIRE(nullptr, ctx, "mom_gpu_horz_advect_tracer: NULL context");
  side_join(ctx);
  REQUIRE(ctx, T_taum1 && uhrho_et && vhrho_nt && th_tendency,
          "mom_gpu_horz_advect_tracer: NULL required array");
  const size_t b3 = p3bytes(ctx);
  const size_t n3 = b3 / sizeof(double);
  int rc0 = MOM_GPU_OK;
  if (wrk1 && (rc0 = launch_fill(ctx, wrk1, 0.0, n3))) return rc0;
  if (flux_x && (rc0 = launch_fill(ctx, flux_x, 0.0, n3))) return rc0;
  if (flux_y && (rc0 = launch_fill(ctx, flux_y, 0.0, n3))) return rc0;
  switch (scheme) {
    case MOM_ADVECT_MDFL_SWEBY:
    case MOM_ADVECT_DST_LINEAR:
      REQUIRE(ctx, wrho_bt && rho_dzt_tau, "horz_advect_tracer(MDFL): wrho_bt/rho_dzt required");
      REQUIRE(ctx, dtime > 0.0, "horz_advect_tracer(MDFL): dtime must be positive");
      return launch_mdfl(ctx, 0, T_taum1, uhrho_et, vhrho_nt, wrho_bt, rho_dzt_tau, dtime,
                         scheme == MOM_ADVECT_MDFL_SWEBY ? 1.0 : 0.0, th_tendency, wrk1, flux_x,
                         flux_y, flux_z);
    case MOM_ADVECT_QUICKER:
      REQUIRE(ctx, T_tau, "horz_advect_tracer(QUICKER): T_tau required");
      REQUIRE(ctx, !limit_with_upwind || tmask_limit,
              "horz_advect_tracer(QUICKER): limit_with_upwind needs tmask_limit");
      return launch_quicker_horz(ctx, T_taum1, T_tau, uhrho_et, vhrho_nt, tmask_limit,
                                 limit_with_upwind, th_tendency, wrk1, flux_x, flux_y);
    default:
      return mom_fail(ctx, MOM_GPU_ERR_ARG,
                      "==>Error from ocean_tracer_advect_mod (horz_advect_tracer): chose invalid "
                      "horz advection scheme (%d)", scheme);
  }
}

int mom_gpu_vert_advect_tracer(mom_gpu_ctx_t *ctx, int scheme, const double *T_taum1,
                               const double *T_tau, const double *wrho_bt,
                               const double *tmask_limit, double *th_tendency, double *wrk1,
                               double *flux_z) {
  REQUIRE(nullptr, ctx, "mom_gpu_vert_advect_tracer: NULL context");
  side_join(ctx);
  REQUIRE(ctx, th_tendency, "mom_gpu_vert_advect_tracer: th_tendency required");
  switch (scheme) {
    case MOM_ADVECT_MDFL_SWEBY:
    case MOM_ADVECT_DST_LINEAR:
      return MOM_GPU_OK;   // three-dimensional schemes: done in horz_advect_tracer
    case MOM_ADVECT_QUICKER:
      REQUIRE(ctx, T_taum1 && T_tau && wrho_bt, "vert_advect_tracer(QUICKER): NULL required array");
      if (wrk1) {
        int rc = launch_fill(ctx, wrk1, 0.0, p3bytes(ctx) / sizeof(double));
        if (rc) return rc;
      }
      return launch_quicker_vert(ctx, T_taum1, T_tau, wrho_bt, tmask_limit, th_tendency, wrk1,
                                 flux_z);
    default:
      return mom_fail(ctx, MOM_GPU_ERR_ARG,
                      "==>Error from ocean_tracer_advect_mod (vert_advect_tracer): invalid "
                      "advection scheme chosen (%d)", scheme);
  }
}

int mom_gpu_advect_sweby_all(mom_gpu_ctx_t *ctx, int ntr, const double *const *T_taum1,
                             const double *uhrho_et, const double *vhrho_nt,
                             const double *wrho_bt, const double *rho_dzt_tau, double dtime,
                             double *const *th_tendency, double *const *wrk1) {
  REQUIRE(nullptr, ctx, "mom_gpu_advect_sweby_all: NULL context");
  side_join(ctx);
  REQUIRE(ctx, ntr >= 0 && T_taum1 && th_tendency && uhrho_et && vhrho_nt && wrho_bt && rho_dzt_tau,
          "mom_gpu_advect_sweby_all: NULL required array");
  REQUIRE(ctx, dtime > 0.0, "mom_gpu_advect_sweby_all: dtime must be positive");
  for (int n = 0; n < ntr; ++n) {
    REQUIRE(ctx, T_taum1[n] && th_tendency[n], "mom_gpu_advect_sweby_all: tracer %d NULL", n);
    double *w1 = wrk1 ? wrk1[n] : nullptr;
    if (w1) {
      int rc0 = launch_fill(ctx, w1, 0.0, p3bytes(ctx) / sizeof(double));
      if (rc0) return rc0;
    }
    int rc = launch_mdfl(ctx, 1, T_taum1[n], uhrho_et, vhrho_nt, wrho_bt, rho_dzt_tau, dtime, 1.0,
                         th_tendency[n], w1, nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  return MOM_GPU_OK;
}

int mom_gpu_compute_tmask_limit(mom_gpu_ctx_t *ctx, int ntr, const double *const *T_tau,
                                const double *min_limit, const double *max_limit, int index_temp,
                                int index_salt, int ts_same, double *const *tmask_limit) {
  REQUIRE(nullptr, ctx, "mom_gpu_compute_tmask_limit: NULL context");
  side_join(ctx);
  REQUIRE(ctx, ntr >= 0 && T_tau && min_limit && max_limit && tmask_limit,
          "mom_gpu_compute_tmask_limit: NULL argument");
  REQUIRE(ctx, index_temp < ntr && index_salt < ntr, "mom_gpu_compute_tmask_limit: bad index");
  for (int n = 0; n < ntr; ++n)
    REQUIRE(ctx, T_tau[n] && tmask_limit[n], "mom_gpu_compute_tmask_limit: tracer %d NULL", n);
  if (ntr == 0) return MOM_GPU_OK;
  return launch_tmask_limit(ctx, ntr, T_tau, min_limit, max_limit, index_temp, index_salt,
                            ts_same, tmask_limit);
}

int mom_gpu_update_tracer_explicit(mom_gpu_ctx_t *ctx, const double *T_taum1,
                                   const double *th_tendency, const double *rho_dzt_taum1,
                                   const double *rho_dztr, double dtime, double *T_taup1) {
  REQUIRE(nullptr, ctx, "mom_gpu_update_tracer_explicit: NULL context");
  side_join(ctx);
  REQUIRE(ctx, T_taum1 && th_tendency && rho_dzt_taum1 && rho_dztr && T_taup1,
          "mom_gpu_update_tracer_explicit: NULL argument");
  return launch_explicit(ctx, T_taum1, th_tendency, rho_dzt_taum1, rho_dztr, dtime, T_taup1);
}

int mom_gpu_invtri(mom_gpu_ctx_t *ctx, double *z, const double *topbc, const double *botbc,
                   const double *dcb, double tdt, const int *kmz, const double *mask,
                   const double *dh, const double *dhw, double aidif) {
  REQUIRE(nullptr, ctx, "mom_gpu_invtri: NULL context");
  side_join(ctx);
  REQUIRE(ctx, z && topbc && botbc && dcb && kmz && mask && dh && dhw, "mom_gpu_invtri: NULL argument");
  return launch_invtri(ctx, z, topbc, botbc, dcb, tdt, kmz, mask, dh, dhw, aidif);
}

int mom_gpu_vert_diffuse_implicit(mom_gpu_ctx_t *ctx, int ntr, double *const *T_taup1,
                                  const double *const *stf, const double *const *btf,
                                  const double *diff_cbt, const double *const *K33_implicit,
                                  int index_salt, double rho0, const double *rho_dzt_taup1,
                                  const double *dzwt, double dtime_t, double aidif) {
  REQUIRE(nullptr, ctx, "mom_gpu_vert_diffuse_implicit: NULL context");
  side_join(ctx);
  REQUIRE(ctx, ntr >= 0 && T_taup1 && stf && btf && diff_cbt && rho_dzt_taup1 && dzwt,
          "mom_gpu_vert_diffuse_implicit: NULL argument");
  for (int n = 0; n < ntr; ++n)
    REQUIRE(ctx, T_taup1[n] && stf[n] && btf[n], "mom_gpu_vert_diffuse_implicit: tracer %d NULL", n);
  if (aidif == 0.0 || ntr == 0) return MOM_GPU_OK;   // ocean_vert_mix.F90:3446
  return launch_vdiff_implicit(ctx, ntr, T_taup1, stf, btf, diff_cbt, K33_implicit, index_salt,
                               rho0, rho_dzt_taup1, dzwt, dtime_t, aidif);
}

int mom_gpu_update_ocean_barotropic(mom_gpu_ctx_t *ctx, const mom_gpu_bt_params_t *p,
                                    const double *eta_t_tau, const double *eta_t_bar_tau,
                                    const double *udrho_tau, const double *forcing_bt,
                                    const double *thicku, const double *patm, const double *pme,
                                    const double *river, const double *source,
                                    const double *eta_smooth, const double *rho_g,
                                    double *eta_t_bar_taup1, double *udrho_taup1,
                                    double *conv_rho_ud_t) {
  REQUIRE(nullptr, ctx, "mom_gpu_update_ocean_barotropic: NULL context");
  REQUIRE(ctx, p, "mom_gpu_update_ocean_barotropic: NULL params");
  REQUIRE(ctx, eta_t_bar_tau && udrho_tau && forcing_bt && thicku && patm && pme && river &&
                   source && eta_smooth && rho_g && eta_t_bar_taup1 && udrho_taup1,
          "mom_gpu_update_ocean_barotropic: NULL argument");
  REQUIRE(ctx, p->initsum_with_bar || eta_t_tau,
          "mom_gpu_update_ocean_barotropic: eta_t_tau required when initsum_with_bar is false");
  REQUIRE(ctx, p->nts >= 1, "mom_gpu_update_ocean_barotropic: nts must be >= 1");
  REQUIRE(ctx, p->dtbt > 0.0 && p->rho0 > 0.0, "mom_gpu_update_ocean_barotropic: bad dtbt/rho0");
  REQUIRE(ctx, !ctx->kd.cyclic_y, "mom_gpu_update_ocean_barotropic: cyclic_y is not supported");
  {
    const void *outs[] = {eta_t_bar_taup1, udrho_taup1, conv_rho_ud_t};
    const void *ins[] = {eta_t_tau, eta_t_bar_tau, udrho_tau};
    for (const void *o : outs)
      for (const void *i : ins)
        REQUIRE(ctx, !o || o != i,
                "mom_gpu_update_ocean_barotropic: an output aliases eta_t_tau, eta_t_bar_tau "
                "or udrho_tau");
  }
  return launch_barotropic(ctx, p, eta_t_tau ? eta_t_tau : eta_t_bar_tau, eta_t_bar_tau, udrho_tau,
                           forcing_bt, thicku, patm, pme, river, source, eta_smooth, rho_g,
                           eta_t_bar_taup1, udrho_taup1, conv_rho_ud_t);
}

static int check_step(mom_gpu_ctx *ctx, const mom_gpu_tracer_step_t *s, const char *who) {
  REQUIRE(ctx, s, "%s: NULL descriptor", who);
  REQUIRE(ctx, s->ntr >= 0 && (s->ntr == 0 || (s->T_taum1 && s->T_taup1 && s->dtype)),
          "%s: NULL tracer tables", who);
  REQUIRE(ctx, s->uhrho_et && s->vhrho_nt && s->wrho_bt && s->rho_dzt_tau,
          "%s: NULL advection field", who);
  REQUIRE(ctx, s->dtime > 0.0, "%s: dtime must be positive", who);
  for (int n = 0; n < s->ntr; ++n) {
    REQUIRE(ctx, s->dtype[n] == 8 || s->dtype[n] == 4, "%s: dtype[%d]=%d", who, n, s->dtype[n]);
    if (s->scheme) {
      REQUIRE(ctx, s->scheme[n] == MOM_ADVECT_MDFL_SWEBY || s->scheme[n] == MOM_ADVECT_QUICKER,
              "%s: tracer %d: scheme %d (MDFL_SWEBY or QUICKER)", who, n, s->scheme[n]);
      REQUIRE(ctx, s->scheme[n] != MOM_ADVECT_QUICKER || s->dtype[n] == 8,
              "%s: tracer %d: Quicker tracers are FP64", who, n);
      REQUIRE(ctx, s->scheme[n] != MOM_ADVECT_QUICKER || (s->min_limit && s->max_limit),
              "%s: tracer %d: Quicker needs min_limit and max_limit", who, n);
    }
    REQUIRE(ctx, s->T_taum1[n] && s->T_taup1[n], "%s: tracer %d NULL", who, n);
    REQUIRE(ctx, s->T_taum1[n] != (const void *)s->T_taup1[n],
            "%s: tracer %d: taum1 and taup1 must differ", who, n);
  }
  if (s->scheme && s->tmask_limit_ts_same) {
    REQUIRE(ctx, s->index_temp < s->ntr && s->index_salt < s->ntr,
            "%s: index_temp/index_salt out of range", who);
    const bool tq = s->index_temp >= 0 && s->scheme[s->index_temp] == MOM_ADVECT_QUICKER;
    const bool sq = s->index_salt >= 0 && s->scheme[s->index_salt] == MOM_ADVECT_QUICKER;
    REQUIRE(ctx, !(tq || sq) || (s->index_temp >= 0 && s->index_salt >= 0 &&
                                 s->dtype[s->index_temp] == 8 && s->dtype[s->index_salt] == 8),
            "%s: tmask_limit_ts_same needs FP64 temperature and salinity", who);
  }
  if (s->stf && s->btf && s->aidif != 0.0) {
    REQUIRE(ctx, s->dzwt && s->diff_cbt, "%s: dzwt/diff_cbt required", who);
    for (int n = 0; n < s->ntr; ++n)
      REQUIRE(ctx, s->stf[n] && s->btf[n], "%s: stf/btf %d NULL", who, n);
  }
  return MOM_GPU_OK;
}

// The tracers of dtype[n] (the others are passed as 0 and skipped) on the
// context's current stream: advection + explicit update, then the implicit
// vertical diffusion.
static int step_part(mom_gpu_ctx *ctx, const mom_gpu_tracer_step_t *s, const int *dtype) {
  const double *rm1 = s->rho_dzt_taum1 ? s->rho_dzt_taum1 : s->rho_dzt_tau;
  // MDFL-Sweby tracers through the grouped passes, Quicker ones one pass each
  std::vector<int> d_mdfl(dtype, dtype + s->ntr), q(s->ntr, 0);
  bool any_q = false;
  if (s->scheme)
    for (int n = 0; n < s->ntr; ++n)
      if (dtype[n] && s->scheme[n] == MOM_ADVECT_QUICKER) {
        d_mdfl[n] = 0;
        q[n] = MOM_ADVECT_QUICKER;
        any_q = true;
      }
  int rc = launch_tracer_step(ctx, s->ntr, s->T_taum1, s->T_taup1, d_mdfl.data(), s->th_tendency,
                              s->uhrho_et, s->vhrho_nt, s->wrho_bt, s->rho_dzt_tau, rm1,
                              s->rho_dztr, s->dtime);
  if (rc) return rc;
  if (any_q) {
    rc = launch_quicker_step(ctx, s->ntr, s->T_taum1, s->T_taup1, q.data(), s->th_tendency,
                             s->min_limit, s->max_limit, s->index_temp, s->index_salt,
                             s->tmask_limit_ts_same, s->limit_with_upwind, s->uhrho_et,
                             s->vhrho_nt, s->wrho_bt, s->rho_dzt_tau, rm1, s->rho_dztr, s->dtime);
    if (rc) return rc;
  }
  if (s->stf && s->btf && s->aidif != 0.0) {
    const double *dh = s->rho_dzt_taup1 ? s->rho_dzt_taup1 : s->rho_dzt_tau;
    rc = launch_vdiff_mixed(ctx, s->ntr, s->T_taup1, dtype, s->stf, s->btf, s->diff_cbt,
                            s->index_salt, s->rho0, dh, s->dzwt, s->dtime, s->aidif);
  }
  return rc;
}

int mom_gpu_step_tracers(mom_gpu_ctx_t *ctx, const mom_gpu_tracer_step_t *s) {
  REQUIRE(nullptr, ctx, "mom_gpu_step_tracers: NULL context");
  int rc = check_step(ctx, s, "mom_gpu_step_tracers");
  if (rc) return rc;
  if (s->ntr == 0) return MOM_GPU_OK;
  // step_streams = 2: the FP32 tracers go on the second stream, where their
  // (VALU-bound) advection pass overlaps the main stream's memory-bound
  // solves and barotropic subcycles.  They read only this call's inputs and
  // write only FP32 arrays, which no main-stream work of a step touches, and
  // use no context scratch (the register-resident solver).
  std::vector<int> d64(s->ntr, 0), d32(s->ntr, 0);
  int n32 = 0;
  for (int n = 0; n < s->ntr; ++n) {
    if (s->dtype[n] == 4) { d32[n] = 4; ++n32; } else { d64[n] = s->dtype[n]; }
  }
  const bool diffuse = s->stf && s->btf && s->aidif != 0.0;
  const bool two = ctx->tune.step_streams == 2 && n32 > 0 &&
                   (!diffuse || !vdiff_fp32_scratch(ctx));
  if (!two) {
    side_join(ctx);
    return step_part(ctx, s, s->dtype);
  }
  // fork: the second stream waits for the main stream's position at the start
  // of the previous step call (only barotropic work may have been queued
  // since; it touches none of a step's arrays), else for its position now
  if (!ctx->fork_valid) MOM_HIP(ctx, hipEventRecord(ctx->fork_ev, ctx->stream));
  MOM_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
  MOM_HIP(ctx, hipEventRecord(ctx->fork_ev, ctx->stream));
  ctx->fork_valid = true;
  if (n32 < s->ntr) {
    rc = step_part(ctx, s, d64.data());
    if (rc) return rc;
  }
  hipStream_t main = ctx->stream;
  ctx->stream = ctx->side;
  rc = step_part(ctx, s, d32.data());
  ctx->stream = main;
  if (rc) return rc;
  MOM_HIP(ctx, hipEventRecord(ctx->side_done, ctx->side));
  ctx->side_pending = true;
  return MOM_GPU_OK;
}

int mom_gpu_tracer_content(mom_gpu_ctx_t *ctx, const double *T, const double *rho_dzt,
                           double *content) {
  REQUIRE(nullptr, ctx, "mom_gpu_tracer_content: NULL context");
  side_join(ctx);
  REQUIRE(ctx, T && rho_dzt && content, "mom_gpu_tracer_content: NULL argument");
  return launch_content(ctx, T, 8, rho_dzt, content);
}

int mom_gpu_tracer_content_f32(mom_gpu_ctx_t *ctx, const float *T, const double *rho_dzt,
                               double *content) {
  REQUIRE(nullptr, ctx, "mom_gpu_tracer_content_f32: NULL context");
  side_join(ctx);
  REQUIRE(ctx, T && rho_dzt && content, "mom_gpu_tracer_content_f32: NULL argument");
  return launch_content(ctx, T, 4, rho_dzt, content);
}

int mom_gpu_tracer_error(mom_gpu_ctx_t *ctx, const void *a, int dtype, const double *b,
                         double rel_floor, double *out) {
  REQUIRE(nullptr, ctx, "mom_gpu_tracer_error: NULL context");
  side_join(ctx);
  REQUIRE(ctx, a && b && out, "mom_gpu_tracer_error: NULL argument");
  REQUIRE(ctx, dtype == 8 || dtype == 4, "mom_gpu_tracer_error: dtype must be 8 or 4");
  return launch_tracer_error(ctx, a, dtype, b, rel_floor, out);
}

}  // extern "C"
