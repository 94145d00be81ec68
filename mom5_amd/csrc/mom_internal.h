// Internal declarations shared by the HIP translation units of libmom5_gpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "mom5_gpu.h"
#include "table_cache.h"

// Compact domain description passed to kernels by value.  Compute-domain
// indices are 0-based (gi in [0, ni), gj in [0, nj)); the data domain has a
// one-cell halo, so the data-domain element of (gi, gj) is
// (gi+1) + NI*(gj+1) and gi = -1 / gi = ni address the halo columns.
struct KDom {
  int ni, nj, nk;
  int NI, NJ;
  int P2;              // NI*NJ, stride of one k level
  int cyclic_x, cyclic_y;
};

__host__ __device__ inline int d2(const KDom &d, int gi, int gj) {
  return (gi + 1) + d.NI * (gj + 1);
}

// Position t of a launch order -> row-major tile index of an ntx x nty grid.
// Tiles are taken in strips S tiles wide, top to bottom within a strip, so
// that a tile's y-neighbour follows it S positions later (not a whole tile
// row) and the halo rows both read are still in L2; the last strip is
// narrower when S does not divide ntx.  A bijection of [0, ntx*nty).
template <int S>
__device__ __forceinline__ int strip_tile(int t, int ntx, int nty) {
  const int full = ntx - ntx % S;   // columns in whole strips
  int row, col;
  if (t < full * nty) {
    const int s = t / (S * nty), r = t - s * (S * nty);
    row = r / S;
    col = s * S + r % S;
  } else {
    const int w = ntx - full, r = t - full * nty;
    row = r / w;
    col = full + r % w;
  }
  return row * ntx + col;
}

// Wrap of a compute index on a cyclic axis.
__host__ __device__ inline int wrapi(int g, int n) {
  g %= n;
  return g < 0 ? g + n : g;
}

// Source point of a halo update on a single PE (mpp_update_domains,
// shared/mpp/include/mpp_do_update.h:186-288): a halo column on a compute
// row takes its cyclic-x image, a halo row on a compute column its cyclic-y
// image, a corner only when both axes are cyclic.  Returns false where the
// update leaves the point alone (the point itself is the source).
__device__ inline bool halo_source(const KDom &d, int ii, int jj, int &si, int &sj) {
  si = ii;
  sj = jj;
  const bool col = ii < 1 || ii > d.ni, row = jj < 1 || jj > d.nj;
  if (!col && !row) return true;
  if (col && row) {
    if (!(d.cyclic_x && d.cyclic_y)) return false;
    si = 1 + wrapi(ii - 1, d.ni);
    sj = 1 + wrapi(jj - 1, d.nj);
    return true;
  }
  if (col) {
    if (!d.cyclic_x) return false;
    si = 1 + wrapi(ii - 1, d.ni);
    return true;
  }
  if (!d.cyclic_y) return false;
  sj = 1 + wrapi(jj - 1, d.nj);
  return true;
}

// Fortran MIN/MAX on non-NaN data.  The ternary form pins which operand
// is returned for +0/-0 so that the result matches the CPU restatement.
// Neighbour lanes of a 64-wide region row through DPP wavefront shifts
// (v_mov_b32_dpp wave_shr:1 / wave_shl:1, full rate, no LDS traffic).  The
// first (last) lane receives 0: callers use lane-1 only from lane 1 up and
// lane+1, lane+2 only up to lane 62, 61.  A lane whose source lane is
// inactive also receives 0, so callers shift outside divergent branches.
__device__ inline int dpp_prev(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xf, 0xf, true); }
__device__ inline int dpp_next(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xf, 0xf, true); }
__device__ inline float lane_prev(float x) {
  return __builtin_bit_cast(float, dpp_prev(__builtin_bit_cast(int, x)));
}
__device__ inline float lane_next(float x) {
  return __builtin_bit_cast(float, dpp_next(__builtin_bit_cast(int, x)));
}
__device__ inline double lane_prev(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned long long lo = (unsigned)dpp_prev((int)(unsigned)b);
  const unsigned long long hi = (unsigned)dpp_prev((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, lo | (hi << 32));
}
__device__ inline double lane_next(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned long long lo = (unsigned)dpp_next((int)(unsigned)b);
  const unsigned long long hi = (unsigned)dpp_next((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, lo | (hi << 32));
}
__device__ inline int lane_prev(int x) { return dpp_prev(x); }
__device__ inline int lane_next(int x) { return dpp_next(x); }

__host__ __device__ inline double fmn(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ inline double fmx(double a, double b) { return (b > a) ? b : a; }
__host__ __device__ inline float fmnf(float a, float b) { return (b < a) ? b : a; }
__host__ __device__ inline float fmxf(float a, float b) { return (b > a) ? b : a; }

struct mom_gpu_ctx {
  mom_gpu_domain_t dom;
  KDom kd;
  int device;
  hipStream_t stream;
  hipEvent_t ev[8];
  // Second stream (default priority): the FP32 passive-tracer work of a
  // batched step when step_streams = 2.  side_done marks the last work
  // queued there; fork_ev the main stream's position at the start of the
  // previous step call, which a step's side work waits for (fork_valid:
  // nothing else touched the step's inputs since).  side_join makes the
  // main stream wait for pending side work.
  hipStream_t side;
  hipEvent_t side_done, fork_ev;
  // ordering of ensemble calls against the member's own stream
  hipEvent_t join_ev;
  bool side_pending, fork_valid;
  // device copies of the grid (owned)
  double *dxt, *dyt, *dat, *datr, *dxte, *dyte, *dxtn, *dytn;
  double *dxu, *dyu, *dxur, *dyur, *f;
  int *kmt, *kmu;
  double *dzt;
  // Quicker weights (quicker_init): 16 data-domain planes
  // quick_x(2) quick_y(2) curv_xp(3) curv_xn(3) curv_yp(3) curv_yn(3),
  // and quick_z(nk,2) curv_zp(nk,3) curv_zn(nk,3).
  double *qw;
  double *qzw;
  // barotropic work: eta_bt(2 levels), udrho_bt(2 comps x 2 levels),
  // steady forcing, cori1, cori2
  double *bt;
  // the grid's dxur, dyur are 1/(dxu + epsln), 1/(dyu + epsln) bit for bit
  // (ocean_grids.F90:949-951): the blocked barotropic forms them, not loads
  int bt_recip;
  // the grid metrics the blocked barotropic multiplies are bounded by 1e100
  // and dxu, dyu, datr carry no sign bit: the condition its land path needs
  int bt_grid_ok;
  double bt_cori_dtbt;
  // REMAP_BT_TO_BU metrics dte, dtw, dus, dun, daur (5 planes, optional)
  double *remap;
  // state of the barotropic variants (18 planes, allocated on first use)
  double *btv;
  // invtri work (e coefficients), one 3-D plane set
  double *tri_e;
  double *tri_e2;   // second factorisation plane (salt's mixing class), lazy
  // FP32 implicit diffusion: forward increments, one plane per tracer of
  // the largest solver group seen (lazy)
  float *vd_scr;
  int vd_scr_planes;
  // reduction partials
  double *red;
  int red_n;
  // launch-shape knobs (mom_gpu_set_tuning)
  struct {
    int bt_subcycles_per_launch;   // 0: one launch per phase per subcycle
    int mdfl_rows_per_wave;        // 2 or 4 region rows per wavefront
    int step_group_f64;            // FP64 tracers per pass of the batched step (1, 2)
    int step_group_f32;            // FP32 tracers per pass (1, 2, 4)
    int step_rows_per_wave;        // 2 or 4 (FP64 groups)
    int step_rows_per_wave_f32;    // 2 or 4 (FP32 groups)
    int step_waves_f32;            // 8 or 16 wavefronts per tile (FP32 groups)
    int step_waves_f64;            // 8 or 12 wavefronts per tile (FP64 groups)
    int vdiff_prefetch;            // FP32 implicit diffusion: loads one level ahead (0, 1)
    int vdiff_block;               // FP32 implicit diffusion: threads a block (64, 256)
    int vdiff_merge;               // mixed steps: all tracers in one diffusion launch (0, 1)
    int vdiff_solver;              // FP32 diffusion: 0 lane per column + scratch, 1 registers
    int step_streams;              // 2: a step's FP32 tracers on the second stream
    int bt_land_tiles;             // blocked barotropic: all-land tiles take the short path (0, 1)
  } tune;
  char err[512];
};

// The main stream waits for the context's pending second-stream work.
void side_join(mom_gpu_ctx *c);
// Whether the FP32 implicit diffusion of this context uses its solver scratch
// (vdiff_solver 0, or nk outside the register solver's range).
bool vdiff_fp32_scratch(const mom_gpu_ctx *c);

// Error handling: every entry point returns through these.
int mom_fail(mom_gpu_ctx *ctx, int code, const char *fmt, ...);

#define MOM_HIP(ctx, call)                                                  \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return mom_fail((ctx), MOM_GPU_ERR_HIP, "%s failed: %s", #call,       \
                      hipGetErrorString(e_));                               \
  } while (0)

#define MOM_LAUNCHED(ctx, name)                                             \
  do {                                                                      \
    hipError_t e_ = hipGetLastError();                                      \
    if (e_ != hipSuccess)                                                   \
      return mom_fail((ctx), MOM_GPU_ERR_HIP, "launch of %s failed: %s",    \
                      (name), hipGetErrorString(e_));                       \
  } while (0)

// Kernel launchers implemented in the k_*.hip translation units.
int launch_mdfl(mom_gpu_ctx *ctx, int flavor, const double *T,
                const double *u, const double *v, const double *w,
                const double *rho, double dtime, double lim, double *th,
                double *wrk1, double *fx, double *fy, double *fz);
int launch_quicker_init(mom_gpu_ctx *ctx);
int launch_quicker_horz(mom_gpu_ctx *ctx, const double *Tm1, const double *Tt,
                        const double *u, const double *v,
                        const double *tlim, int limit_with_upwind,
                        double *th, double *wrk1, double *fx, double *fy);
int launch_quicker_vert(mom_gpu_ctx *ctx, const double *Tm1, const double *Tt,
                        const double *w, const double *tlim, double *th,
                        double *wrk1, double *fz);
int launch_vdiff_mixed(mom_gpu_ctx *ctx, int ntr, void *const *T, const int *dtype,
                       const double *const *stf, const double *const *btf,
                       const double *diff_cbt, int index_salt, double rho0, const double *dh,
                       const double *dhw, double dt, double aidif);
int launch_tracer_step(mom_gpu_ctx *ctx, int ntr, const void *const *T_in, void *const *T_out,
                       const int *dtype, const double *const *th_in, const double *u,
                       const double *v, const double *w, const double *rho,
                       const double *rho_taum1, const double *rho_dztr, double dtime);
int launch_quicker_step(mom_gpu_ctx *ctx, int ntr, const void *const *T_in, void *const *T_out,
                        const int *scheme, const double *const *th_in, const double *mn,
                        const double *mx, int it, int is, int ts_same, int lwu,
                        const double *u, const double *v, const double *w, const double *rho,
                        const double *rho_taum1, const double *rho_dztr, double dtime);
int launch_tmask_limit(mom_gpu_ctx *ctx, int ntr, const double *const *T,
                       const double *mn, const double *mx, int it, int is,
                       int ts_same, double *const *lim);
int launch_explicit(mom_gpu_ctx *ctx, const double *Tm1, const double *th,
                    const double *rdz, const double *rdzr, double dtime,
                    double *Tp1);
int launch_invtri(mom_gpu_ctx *ctx, double *z, const double *topbc,
                  const double *botbc, const double *dcb, double tdt,
                  const int *kmz, const double *mask, const double *dh,
                  const double *dhw, double aidif);
int launch_vdiff_implicit(mom_gpu_ctx *ctx, int ntr, double *const *T,
                          const double *const *stf, const double *const *btf,
                          const double *diff_cbt,
                          const double *const *K33, int index_salt,
                          double rho0, const double *dh, const double *dhw,
                          double dt, double aidif);
int launch_barotropic(mom_gpu_ctx *ctx, const mom_gpu_bt_params_t *p,
                      const double *eta_t, const double *eta_bar_tau,
                      const double *ud_tau, const double *forcing,
                      const double *thicku, const double *patm,
                      const double *pme, const double *river,
                      const double *source, const double *eta_smooth,
                      const double *rho_g, double *eta_bar, double *ud,
                      double *conv);
int launch_barotropic_blocked(mom_gpu_ctx *ctx, const mom_gpu_bt_params_t *p,
                              const double *eta_t, const double *eta_bar_tau,
                              const double *ud_tau, const double *forcing, const double *thicku,
                              const double *patm, const double *pme, const double *river,
                              const double *source, const double *eta_smooth,
                              const double *rho_g, double *eta_bar, double *ud, double *conv,
                              int hsub);
int launch_content(mom_gpu_ctx *ctx, const void *T, int dtype, const double *rho, double *out);
// the content kernels with their partial sums (2 doubles per block, ctx->red_n
// blocks) in `part` and the result in device memory `out_dev`; no copy, no sync
int launch_content_part(mom_gpu_ctx *ctx, const void *T, int dtype, const double *rho,
                        double *part, double *out_dev);
int launch_tracer_error(mom_gpu_ctx *ctx, const void *a, int dtype, const double *b,
                        double floor, double *out);
int launch_fill(mom_gpu_ctx *ctx, double *p, double v, size_t n);
int launch_update_domains(mom_gpu_ctx *ctx, double *a, int nz);
int launch_bt_forcing(mom_gpu_ctx *ctx, const double *accel, const double *smf, const double *bmf,
                      double *forcing_bt);
int launch_bt_variant(mom_gpu_ctx *ctx, const mom_gpu_bt_params_t *p,
                      const mom_gpu_bt_variant_t *v);
int launch_eta_update(mom_gpu_ctx *ctx, int pressure_based, double rho0, double grav,
                      double dtime, const double *conv, const double *pme, const double *river,
                      const double *source, const double *dpatm_dt, const double *ht,
                      const double *x_taum1, double *x_taup1, double *pbot, double *tend);
int launch_bt_closure(mom_gpu_ctx *ctx, int pressure_based, double rho0, double grav,
                      const double *eta_bar, const double *anompb_bar, const double *patm,
                      const double *rho_g, const double *ht, const double *tmasken,
                      double *eta_u, double *ps, double *grad_ps, double *grad_ab);
int launch_adv_vel(mom_gpu_ctx *ctx, const double *u, const double *rho_dzu, const double *pme,
                   const double *river, const double *rtend, const double *msrc, double *ue,
                   double *vn, double *wb, double *div);
int launch_copy(mom_gpu_ctx *ctx, double *dst, const double *src, size_t n);
int launch_check_recip(mom_gpu_ctx *ctx);
int launch_check_masks(mom_gpu_ctx *ctx, const double *tmask,
                       const double *umask, long long *bad);

// ---- ensembles of domains ---------------------------------------------------
struct mom_gpu_ensemble {
  std::vector<mom_gpu_ctx *> ctx;   // not owned; ctx[0] is the lead (stream, errors)
  mom_gpu_ctx *lead = nullptr;
  int *counters = nullptr;          // device work-queue counters
  static constexpr int NCOUNT = 16;
  TableCache<void *> tables;        // launch tables, shared by content
};

// Device copy of a launch table; identical tables (the two parities of
// two-level time stepping) share one copy.  Uploading a new table is refused
// while the lead stream is being captured; tables used during a capture are
// kept until the ensemble is destroyed.  nullptr on failure.
const void *ens_table(mom_gpu_ensemble *e, const void *host, size_t bytes);
int launch_step_ensemble(mom_gpu_ensemble *e, const mom_gpu_tracer_step_t *steps);
int launch_vdiff_ensemble(mom_gpu_ensemble *e, const mom_gpu_tracer_step_t *steps);
int launch_barotropic_ensemble(mom_gpu_ensemble *e, const mom_gpu_bt_params_t *p,
                               const mom_gpu_bt_fields_t *f);
int launch_set_ints(mom_gpu_ctx *ctx, int *p, int v, int n);
