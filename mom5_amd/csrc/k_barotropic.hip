// Split-explicit barotropic subcycle, B-grid, depth-based vertical
// coordinate: pred_corr_tropic_depth_bgrid (ocean_barotropic.F90:3225-3579)
// with barotropic_halo=1 and use_legacy_barotropic_halos=.true., and the
// time-averaging tail of update_ocean_barotropic (:2804-2811, :2846-2888).
//
// Everything is computed on the data domain, as the reference does.  The one
// halo exchange inside the loop (eta_t_bt after the predictor, :3383) is
// folded into the predictor: on a cyclic-x domain the halo columns of the
// compute rows take the value computed for the interior column they mirror.
#include "mom_internal.h"

namespace {

struct BtBuf {
  double *eta;    // eta_t_bt, 2 levels
  double *ud;     // udrho_bt, (2 components) x (2 levels)
  double *sf;     // steady_forcing
  double *cori1, *cori2;
};

__device__ inline double *ETA(const BtBuf &b, int P2, int lev) { return b.eta + (size_t)P2 * (lev - 1); }
__device__ inline double *UD(const BtBuf &b, int P2, int c, int lev) {
  return b.ud + (size_t)P2 * ((c - 1) + 2 * (lev - 1));
}

// DIV_UD (ocean_operators.F90:507-608, B-grid, non-legacy order) at the
// data-domain point (ii, jj); zero on the first row and column.
__device__ inline double div_ud(const KDom &d, const double *u1, const double *u2,
                                const double *dxu, const double *dyu, const double *datr,
                                int ii, int jj) {
  if (ii == 0 || jj == 0) return 0.0;
  const int o = ii + d.NI * jj, ojm = o - d.NI, oim = o - 1, oimjm = ojm - 1;
  const double uh = u1[o] * dyu[o], vh = u2[o] * dxu[o];
  const double uhjm = u1[ojm] * dyu[ojm], vhjm = u2[ojm] * dxu[ojm];
  const double uhim = u1[oim] * dyu[oim], vhim = u2[oim] * dxu[oim];
  const double uhimjm = u1[oimjm] * dyu[oimjm], vhimjm = u2[oimjm] * dxu[oimjm];
  const double uh_bay = 0.5 * (uh + uhjm);
  const double vh_bax = 0.5 * (vh + vhim);
  const double uhim_bay = 0.5 * (uhim + uhimjm);
  const double vhjm_bax = 0.5 * (vhjm + vhimjm);
  return ((uh_bay - uhim_bay) + (vh_bax - vhjm_bax)) * datr[o];
}

__global__ void k_bt_init(KDom d, BtBuf b, int initsum, double dtbt,
                          const double *__restrict__ eta_t, const double *__restrict__ eta_bar_tau,
                          const double *__restrict__ ud_tau, const double *__restrict__ pme,
                          const double *__restrict__ river, const double *__restrict__ source,
                          const double *__restrict__ eta_smooth, const double *__restrict__ f,
                          const int *__restrict__ kmt, double *__restrict__ eta_bar,
                          double *__restrict__ ud) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj, P2 = d.P2;
  eta_bar[o] = initsum ? eta_bar_tau[o] : eta_t[o];
  ud[o] = 0.0;
  ud[o + P2] = 0.0;
  // itime = 1: fstau = mod(1,2)+1 = 2
  ETA(b, P2, 2)[o] = eta_bar_tau[o];
  UD(b, P2, 1, 2)[o] = ud_tau[o];
  UD(b, P2, 2, 2)[o] = ud_tau[o + P2];
  const double tm1 = (1 <= kmt[o]) ? 1.0 : 0.0;
  b.sf[o] = tm1 * (pme[o] + river[o] + source[o] - eta_smooth[o]);
  const double c1 = 0.5 * dtbt * f[o];       // ocean_barotropic.F90:1285-1286
  b.cori1[o] = c1;
  b.cori2[o] = 1.0 / (1.0 + c1 * c1);
}

// predictor for eta (:3365-3376) with the halo update (:3383) folded in
__global__ void k_bt_pred(KDom d, BtBuf b, int fstau, int fstaup1, double c_pred,
                          const double *__restrict__ dxu, const double *__restrict__ dyu,
                          const double *__restrict__ datr, const int *__restrict__ kmt) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int P2 = d.P2;
  int si = ii;
  if (d.cyclic_x && jj >= 1 && jj <= d.NJ - 2) {
    if (ii == 0) si = d.NI - 2;
    else if (ii == d.NI - 1) si = 1;
  }
  const int os = si + d.NI * jj;
  const double tm1 = (1 <= kmt[os]) ? 1.0 : 0.0;
  const double tmp = -div_ud(d, UD(b, P2, 1, fstau), UD(b, P2, 2, fstau), dxu, dyu, datr, si, jj) * tm1;
  ETA(b, P2, fstaup1)[ii + d.NI * jj] = ETA(b, P2, fstau)[os] + c_pred * (tmp + b.sf[os]);
}

// surface pressure, its B-grid gradient (GRAD_BAROTROPIC_P,
// ocean_operators.F90:631-676) and the momentum update with implicit
// Coriolis (:3394-3486)
__global__ void k_bt_mom(KDom d, BtBuf b, int fstau, int fstaup1, double dtbt,
                         const double *__restrict__ rho_g, const double *__restrict__ patm,
                         const double *__restrict__ thicku, const double *__restrict__ forcing,
                         const double *__restrict__ f, const double *__restrict__ dxur,
                         const double *__restrict__ dyur, const int *__restrict__ kmu) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj, P2 = d.P2;
  const double *eta = ETA(b, P2, fstaup1);
  double g1 = 0.0, g2 = 0.0;
  if (ii < d.NI - 1 && jj < d.NJ - 1) {
    const int oe = o + 1, on = o + d.NI, one = on + 1;
    const double p = rho_g[o] * eta[o] + patm[o];
    const double pe = rho_g[oe] * eta[oe] + patm[oe];
    const double pn = rho_g[on] * eta[on] + patm[on];
    const double pne = rho_g[one] * eta[one] + patm[one];
    const double p_fay = 0.5 * (p + pn);
    const double p_fax = 0.5 * (p + pe);
    const double pip_fay = 0.5 * (pe + pne);
    const double pjp_fax = 0.5 * (pn + pne);
    const double um = (1 <= kmu[o]) ? 1.0 : 0.0;
    g1 = (pip_fay - p_fay) * dxur[o] * um;
    g2 = (pjp_fax - p_fax) * dyur[o] * um;
  }
  const double pf1 = -(thicku[o] * g1);
  const double pf2 = -(thicku[o] * g2);
  const double u1 = UD(b, P2, 1, fstau)[o], u2 = UD(b, P2, 2, fstau)[o];
  const double t1 = u1 + dtbt * (0.5 * f[o] * u2 + pf1 + forcing[o] + 0.0);
  const double t2 = u2 + dtbt * (-(0.5 * f[o] * u1) + pf2 + forcing[o + P2] + 0.0);
  UD(b, P2, 1, fstaup1)[o] = b.cori2[o] * (t1 + b.cori1[o] * t2);
  UD(b, P2, 2, fstaup1)[o] = b.cori2[o] * (t2 - b.cori1[o] * t1);
}

// corrector (:3497-3505) and the accumulation for the time average
// (:3532-3549)
__global__ void k_bt_corr(KDom d, BtBuf b, int fstau, int fstaup1, double c_corr, int step_a,
                          double wgt, const double *__restrict__ dxu,
                          const double *__restrict__ dyu, const double *__restrict__ datr,
                          const int *__restrict__ kmt, double *__restrict__ eta_bar,
                          double *__restrict__ ud) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj, P2 = d.P2;
  const double *u1 = UD(b, P2, 1, fstaup1), *u2 = UD(b, P2, 2, fstaup1);
  const double tm1 = (1 <= kmt[o]) ? 1.0 : 0.0;
  const double tmp = -div_ud(d, u1, u2, dxu, dyu, datr, ii, jj) * tm1;
  const double e = ETA(b, P2, fstau)[o] + c_corr * (tmp + b.sf[o]);
  ETA(b, P2, fstaup1)[o] = e;
  eta_bar[o] = eta_bar[o] + e;
  if (step_a) {
    ud[o] = ud[o] + u1[o];
    ud[o + P2] = ud[o + P2] + u2[o];
  } else {
    ud[o] = ud[o] + wgt * u1[o];
    ud[o + P2] = ud[o + P2] + wgt * u2[o];
  }
}

__global__ void k_bt_scale(KDom d, double rntsp1, double rnts, double *__restrict__ eta_bar,
                           double *__restrict__ ud) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  eta_bar[o] = rntsp1 * eta_bar[o];
  ud[o] = rnts * ud[o];
  ud[o + d.P2] = rnts * ud[o + d.P2];
}

// mpp_update_domains on a cyclic-x single domain: E/W halo columns of the
// compute rows (ncomp planes)
__global__ void k_bt_wrap(KDom d, double *__restrict__ a, int ncomp) {
  const int jj = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (jj > d.NJ - 2) return;
  for (int c = 0; c < ncomp; ++c) {
    double *p = a + (size_t)d.P2 * c + (size_t)d.NI * jj;
    p[0] = p[d.NI - 2];
    p[d.NI - 1] = p[1];
  }
}

__global__ void k_bt_conv(KDom d, const double *__restrict__ ud, const double *__restrict__ dxu,
                          const double *__restrict__ dyu, const double *__restrict__ datr,
                          double *__restrict__ conv) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  conv[ii + d.NI * jj] = -div_ud(d, ud, ud + d.P2, dxu, dyu, datr, ii, jj);
}

}  // namespace

int launch_barotropic(mom_gpu_ctx *ctx, const mom_gpu_bt_params_t *p, const double *eta_t,
                      const double *eta_bar_tau, const double *ud_tau, const double *forcing,
                      const double *thicku, const double *patm, const double *pme,
                      const double *river, const double *source, const double *eta_smooth,
                      const double *rho_g, double *eta_bar, double *ud, double *conv) {
  const KDom &d = ctx->kd;
  const size_t P2 = (size_t)d.P2;
  BtBuf b;
  b.eta = ctx->bt;
  b.ud = ctx->bt + 2 * P2;
  b.sf = ctx->bt + 6 * P2;
  b.cori1 = ctx->bt + 7 * P2;
  b.cori2 = ctx->bt + 8 * P2;
  const double rho0r = 1.0 / p->rho0;
  const double c_pred = p->dtbt * p->pred_corr_gamma * rho0r;   // dtbt_gamma_rho0r, :1196
  const double c_corr = p->dtbt * rho0r;                        // dtbt_rho0r, :1197
  const dim3 grid((d.NI + 255) / 256, d.NJ), block(256);
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_bt_init, grid, block, 0, s, d, b, p->initsum_with_bar, p->dtbt, eta_t,
                     eta_bar_tau, ud_tau, pme, river, source, eta_smooth, ctx->f, ctx->kmt,
                     eta_bar, ud);
  MOM_LAUNCHED(ctx, "k_bt_init");
  for (int itime = 1; itime <= p->nts; ++itime) {
    const int fstau = (itime + 0) % 2 + 1;
    const int fstaup1 = (itime + 1) % 2 + 1;
    hipLaunchKernelGGL(k_bt_pred, grid, block, 0, s, d, b, fstau, fstaup1, c_pred, ctx->dxu,
                       ctx->dyu, ctx->datr, ctx->kmt);
    hipLaunchKernelGGL(k_bt_mom, grid, block, 0, s, d, b, fstau, fstaup1, p->dtbt, rho_g, patm,
                       thicku, forcing, ctx->f, ctx->dxur, ctx->dyur, ctx->kmu);
    const double wgt = (double)(p->nts - itime + 1);
    hipLaunchKernelGGL(k_bt_corr, grid, block, 0, s, d, b, fstau, fstaup1, c_corr,
                       p->stepping_a, wgt, ctx->dxu, ctx->dyu, ctx->datr, ctx->kmt, eta_bar, ud);
  }
  MOM_LAUNCHED(ctx, "k_bt_subcycle");
  const double nts = (double)p->nts;
  const double rntsp1 = 1.0 / (nts + 1.0);
  const double rnts = p->stepping_a ? 1.0 / nts : 2.0 / (nts * (double)(p->nts + 1));
  hipLaunchKernelGGL(k_bt_scale, grid, block, 0, s, d, rntsp1, rnts, eta_bar, ud);
  const dim3 wg((d.NJ + 255) / 256), wb(256);
  if (d.cyclic_x) {
    hipLaunchKernelGGL(k_bt_wrap, wg, wb, 0, s, d, eta_bar, 1);
    hipLaunchKernelGGL(k_bt_wrap, wg, wb, 0, s, d, ud, 2);
  }
  if (conv) {
    hipLaunchKernelGGL(k_bt_conv, grid, block, 0, s, d, ud, ctx->dxu, ctx->dyu, ctx->datr, conv);
    if (d.cyclic_x) hipLaunchKernelGGL(k_bt_wrap, wg, wb, 0, s, d, conv, 1);
  }
  MOM_LAUNCHED(ctx, "k_bt_tail");
  return MOM_GPU_OK;
}
