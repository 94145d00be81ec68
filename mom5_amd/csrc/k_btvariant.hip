// The four split-explicit barotropic schemes of update_ocean_barotropic
// (ocean_barotropic.F90:2829-2843) with barotropic_halo = 1 and no tides,
// geoid, friction, smoothing, blobs or OBC, one template per scheme:
//   <CG=false, PR=false> pred_corr_tropic_depth_bgrid :3225-3579
//   <CG=true,  PR=false> pred_corr_tropic_depth_cgrid :3601-3965
//   <CG=false, PR=true>  pred_corr_tropic_press_bgrid :3989-4329
//   <CG=true,  PR=true>  pred_corr_tropic_press_cgrid :4355-4693
// followed by the tail of update_ocean_barotropic (:2846-2894).
//
// Two launches per subcycle, each over the data domain:
//   k_btv_velocity  udrho_bt(fstaup1) at every point.  The predicted x
//                   (eta_t_bt or anompb_bt) is recomputed at the points the
//                   pressure gradient needs instead of being stored; the
//                   B-grid's legacy halo update of the prediction and the
//                   C-grid's halo update of udrho_bt are folded in by
//                   computing a rewritten halo point at its cyclic source.
//                   The C-grid Adams-Bashforth Coriolis force of this
//                   subcycle is written (compute points only) into the one
//                   of its three levels the subcycle does not read.
//   k_btv_corrector x_bt(fstaup1) from the new transports, and the time sums
//                   (kept in scratch so that the tail can read a halo
//                   point's source while other threads write the outputs).
// Every value is formed with the reference's operations in its order
// (-ffp-contract=off), so the results equal the CPU restatement bit for bit.
#include "mom_internal.h"

namespace {

constexpr double kEpsln = 1.0e-40;   // shared/constants/constants.F90:123

template <bool CG, bool PR>
struct BtvArgs {
  KDom d;
  // grid (context copies)
  const double *dxu, *dyu, *dxur, *dyur, *dat, *datr, *dxte, *dytn, *f;
  const int *kmt, *kmu;
  // inputs
  const double *x_tau, *xbar_tau, *ud_tau, *forcing, *thick, *tmasken, *patm, *rho_g,
      *dpatm_dt, *pme, *river, *source, *smooth;
  // state: x_bt (2 levels), ud_bt (2 comps x 2 levels), sf, cori1, cori2,
  // coriolis force (2 comps x 3 levels), time sums (x, ud 2 comps)
  double *xb, *ud, *sf, *cori1, *cori2, *cf, *accx, *accu;
  double dtbt, grav, rho0r;
  double c_pred, c_corr;       // predictor / corrector factors of the scheme
  int predict;                 // 0: eta_t_bt(fstaup1) = eta_t_bt(fstau)
  int initsum;
};

__device__ inline double tm1(const int *kmt, int o) { return 1 <= kmt[o] ? 1.0 : 0.0; }

// DIV_UD (ocean_operators.F90:507-608), halo 1; zero on the first row and
// column of the data domain.
template <bool CG, bool PR>
__device__ inline double div_ud(const BtvArgs<CG, PR> &a, const double *u1, const double *u2,
                                int ii, int jj) {
  if (ii == 0 || jj == 0) return 0.0;
  const KDom &d = a.d;
  const int o = ii + d.NI * jj, ojm = o - d.NI, oim = o - 1;
  if (CG) {   // :526-536
    const double uhim = u1[oim] * a.dat[oim] * (1.0 / (a.dxte[oim] + kEpsln));
    const double vhjm = u2[ojm] * a.dat[ojm] * (1.0 / (a.dytn[ojm] + kEpsln));
    const double uh = u1[o] * a.dat[o] * (1.0 / (a.dxte[o] + kEpsln));
    const double vh = u2[o] * a.dat[o] * (1.0 / (a.dytn[o] + kEpsln));
    return ((uh - uhim) + (vh - vhjm)) * a.datr[o];
  }
  const int oimjm = ojm - 1;  // B-grid, :570-605
  const double uh = u1[o] * a.dyu[o], vh = u2[o] * a.dxu[o];
  const double uhjm = u1[ojm] * a.dyu[ojm], vhjm = u2[ojm] * a.dxu[ojm];
  const double uhim = u1[oim] * a.dyu[oim], vhim = u2[oim] * a.dxu[oim];
  const double uhimjm = u1[oimjm] * a.dyu[oimjm], vhimjm = u2[oimjm] * a.dxu[oimjm];
  const double uh_bay = 0.5 * (uh + uhjm);
  const double vh_bax = 0.5 * (vh + vhim);
  const double uhim_bay = 0.5 * (uhim + uhimjm);
  const double vhjm_bax = 0.5 * (vhjm + vhimjm);
  return ((uh_bay - uhim_bay) + (vh_bax - vhjm_bax)) * a.datr[o];
}

// x_bt(fstaup1) after the predictor at data point q (no halo semantics).
template <bool CG, bool PR>
__device__ inline double predicted(const BtvArgs<CG, PR> &a, const double *x0, const double *u1,
                                   const double *u2, int qi, int qj) {
  const int o = qi + a.d.NI * qj;
  if (!a.predict) return x0[o];
  const double dv = div_ud(a, u1, u2, qi, qj);
  const double tmp = PR ? -a.grav * dv : -dv * tm1(a.kmt, o);
  return x0[o] + a.c_pred * (tmp + a.sf[o]);
}

// The field whose gradient drives the transports, at data point q:
// ps_bt = rho_g*eta_t_bt + patm (depth), tmask*anompb_bt (pressure).  The
// B-grid reads the prediction after its legacy halo update (:3383, :4140).
template <bool CG, bool PR>
__device__ inline double drive_p(const BtvArgs<CG, PR> &a, const double *x0, const double *u1,
                                 const double *u2, int qi, int qj) {
  int si = qi, sj = qj;
  if (!CG) halo_source(a.d, qi, qj, si, sj);
  const double xp = predicted(a, x0, u1, u2, si, sj);
  const int o = qi + a.d.NI * qj;
  return PR ? tm1(a.kmt, o) * xp : a.rho_g[o] * xp + a.patm[o];
}

// C-grid Coriolis force at compute point o from the transports u1, u2
// (:3711-3727 and :3838-3851).
__device__ inline void cori_force(const KDom &d, const double *f, const double *u1,
                                  const double *u2, int o, double &F1, double &F2) {
  const int oe = o + 1, os = o - d.NI, ose = os + 1, ow = o - 1, on = o + d.NI, onw = on - 1;
  F1 = 0.25 * (f[o] * (u2[o] + u2[oe]) + f[os] * (u2[os] + u2[ose]));
  F2 = -(0.25 * (f[o] * (u1[o] + u1[on]) + f[ow] * (u1[ow] + u1[onw])));
}

template <bool CG, bool PR>
__global__ void k_btv_init(BtvArgs<CG, PR> a) {
  const KDom &d = a.d;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj, P2 = d.P2;
  // itime = 1: fstau = 2, fstaup1 = 1
  a.accx[o] = a.initsum ? a.xbar_tau[o] : a.x_tau[o];
  a.accu[o] = 0.0;
  a.accu[o + P2] = 0.0;
  a.xb[P2 + o] = a.xbar_tau[o];
  a.xb[o] = a.xbar_tau[o];
  a.ud[2 * P2 + o] = a.ud_tau[o];
  a.ud[3 * P2 + o] = a.ud_tau[o + P2];
  const double tm = tm1(a.kmt, o);
  const double s = a.pme[o] + a.river[o] + a.source[o] - a.smooth[o];
  a.sf[o] = PR ? tm * (a.grav * s + a.dpatm_dt[o]) : tm * s;
  if (!CG) {
    const double c1 = 0.5 * a.dtbt * a.f[o];      // :1285-1286
    a.cori1[o] = c1;
    a.cori2[o] = 1.0 / (1.0 + c1 * c1);
  } else {
    double F1 = 0.0, F2 = 0.0;
    if (ii >= 1 && ii <= d.ni && jj >= 1 && jj <= d.nj)
      cori_force(d, a.f, a.ud_tau, a.ud_tau + P2, o, F1, F2);
    for (int lev = 0; lev < 3; ++lev) {
      a.cf[(size_t)P2 * (2 * lev) + o] = F1;
      a.cf[(size_t)P2 * (2 * lev + 1) + o] = F2;
    }
  }
}

// Subcycle itime: levels fstau/fstaup1 (1-based) and the Adams-Bashforth
// levels m0 (written), m1, m2.
template <bool CG, bool PR>
__global__ void k_btv_velocity(BtvArgs<CG, PR> a, int fstau, int fstaup1, int m0, int m1,
                               int m2) {
  const KDom &d = a.d;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const size_t P2 = (size_t)d.P2;
  const double *x0 = a.xb + P2 * (fstau - 1);
  const double *u1 = a.ud + P2 * (2 * (fstau - 1));
  const double *u2 = u1 + P2;
  double *n1 = a.ud + P2 * (2 * (fstaup1 - 1));
  double *n2 = n1 + P2;
  const int o = ii + d.NI * jj;
  // the point whose transports this thread computes (C-grid: the halo
  // update of udrho_bt(fstaup1), :3895-3897)
  int si = ii, sj = jj;
  if (CG) halo_source(d, ii, jj, si, sj);
  const int os = si + d.NI * sj;
  double g1 = 0.0, g2 = 0.0;
  if (si < d.NI - 1 && sj < d.NJ - 1) {
    const double pc = drive_p(a, x0, u1, u2, si, sj);
    const double pe = drive_p(a, x0, u1, u2, si + 1, sj);
    const double pn = drive_p(a, x0, u1, u2, si, sj + 1);
    if (CG) {   // GRAD_BAROTROPIC_P, C-grid (ocean_operators.F90:664-669)
      g1 = (pe - pc) * (1.0 / (a.dxte[os] + kEpsln)) * a.tmasken[os];
      g2 = (pn - pc) * (1.0 / (a.dytn[os] + kEpsln)) * a.tmasken[P2 + os];
    } else {    // B-grid (:651-660)
      const double pne = drive_p(a, x0, u1, u2, si + 1, sj + 1);
      const double p_fay = 0.5 * (pc + pn);
      const double p_fax = 0.5 * (pc + pe);
      const double pip_fay = 0.5 * (pe + pne);
      const double pjp_fax = 0.5 * (pn + pne);
      const double um = 1 <= a.kmu[os] ? 1.0 : 0.0;
      g1 = (pip_fay - p_fay) * a.dxur[os] * um;
      g2 = (pjp_fax - p_fax) * a.dyur[os] * um;
    }
  }
  const double th1 = a.thick[os], th2 = CG ? a.thick[P2 + os] : th1;
  const double pf1 = PR ? -(a.rho0r * th1 * g1) : -(th1 * g1);
  const double pf2 = PR ? -(a.rho0r * th2 * g2) : -(th2 * g2);
  const double v1 = u1[os], v2 = u2[os];
  double r1, r2;
  if (CG) {
    double acc1 = 0.0, acc2 = 0.0;
    if (si >= 1 && si <= d.ni && sj >= 1 && sj <= d.nj) {
      double F1, F2;
      cori_force(d, a.f, u1, u2, os, F1, F2);
      const double *c1 = a.cf + P2 * (2 * (m1 - 1)), *c2 = a.cf + P2 * (2 * (m2 - 1));
      constexpr double ab0 = 23.0 / 12.0, ab1 = -16.0 / 12.0, ab2 = 5.0 / 12.0;  // :617-619
      acc1 = ab0 * F1 + ab1 * c1[os] + ab2 * c2[os];
      acc2 = ab0 * F2 + ab1 * c1[P2 + os] + ab2 * c2[P2 + os];
      if (os == o) {
        double *c0 = a.cf + P2 * (2 * (m0 - 1));
        c0[o] = F1;
        c0[P2 + o] = F2;
      }
    }
    const double s1 = acc1 + pf1 + a.forcing[os];
    const double s2 = acc2 + pf2 + a.forcing[P2 + os];
    r1 = v1 + a.dtbt * a.tmasken[os] * (PR ? s1 : s1 + 0.0);
    r2 = v2 + a.dtbt * a.tmasken[P2 + os] * (PR ? s2 : s2 + 0.0);
  } else {
    const double fo = a.f[os];
    double t1, t2;
    if (PR) {
      t1 = v1 + a.dtbt * (0.5 * fo * v2 + pf1 + a.forcing[os]);
      t2 = v2 + a.dtbt * (-(0.5 * fo * v1) + pf2 + a.forcing[P2 + os]);
    } else {
      t1 = v1 + a.dtbt * (0.5 * fo * v2 + pf1 + a.forcing[os] + 0.0);
      t2 = v2 + a.dtbt * (-(0.5 * fo * v1) + pf2 + a.forcing[P2 + os] + 0.0);
    }
    r1 = a.cori2[os] * (t1 + a.cori1[os] * t2);
    r2 = a.cori2[os] * (t2 - a.cori1[os] * t1);
  }
  n1[o] = r1;
  n2[o] = r2;
}

template <bool CG, bool PR>
__global__ void k_btv_corrector(BtvArgs<CG, PR> a, int fstau, int fstaup1, double wgt) {
  const KDom &d = a.d;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const size_t P2 = (size_t)d.P2;
  const double *x0 = a.xb + P2 * (fstau - 1);
  double *x1 = a.xb + P2 * (fstaup1 - 1);
  const double *n1 = a.ud + P2 * (2 * (fstaup1 - 1));
  const double *n2 = n1 + P2;
  const int o = ii + d.NI * jj;
  const double dv = div_ud(a, n1, n2, ii, jj);
  const double tmp = PR ? -a.grav * dv : -dv * tm1(a.kmt, o);
  const double xn = x0[o] + a.c_corr * (tmp + a.sf[o]);
  x1[o] = xn;
  a.accx[o] = a.accx[o] + xn;
  if (wgt == 0.0) {           // barotropic_time_stepping_A
    a.accu[o] = a.accu[o] + n1[o];
    a.accu[P2 + o] = a.accu[P2 + o] + n2[o];
  } else {
    a.accu[o] = a.accu[o] + wgt * n1[o];
    a.accu[P2 + o] = a.accu[P2 + o] + wgt * n2[o];
  }
}

// Tail (:2846-2894): the time means, each halo-updated, and
// conv_rho_ud_t = -DIV_UD(udrho(taup1)) halo-updated.
template <bool CG, bool PR>
__device__ inline double mean_ud(const BtvArgs<CG, PR> &a, double rnts, int c, int qi, int qj) {
  int si, sj;
  halo_source(a.d, qi, qj, si, sj);
  return rnts * a.accu[(size_t)a.d.P2 * c + si + a.d.NI * sj];
}

template <bool CG, bool PR>
__global__ void k_btv_tail(BtvArgs<CG, PR> a, double rntsp1, double rnts, double *__restrict__ xbar,
                           double *__restrict__ udrho, double *__restrict__ conv) {
  const KDom &d = a.d;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const size_t P2 = (size_t)d.P2;
  const int o = ii + d.NI * jj;
  int si, sj;
  halo_source(d, ii, jj, si, sj);
  const int os = si + d.NI * sj;
  xbar[o] = rntsp1 * a.accx[os];
  udrho[o] = rnts * a.accu[os];
  udrho[P2 + o] = rnts * a.accu[P2 + os];
  if (!conv) return;
  // -DIV_UD of the halo-updated mean at the source point
  double dv = 0.0;
  if (si != 0 && sj != 0) {
    const int oim = os - 1, ojm = os - d.NI;
    const double u = mean_ud(a, rnts, 0, si, sj), v = mean_ud(a, rnts, 1, si, sj);
    const double uim = mean_ud(a, rnts, 0, si - 1, sj);
    const double vjm = mean_ud(a, rnts, 1, si, sj - 1);
    if (CG) {
      const double uhim = uim * a.dat[oim] * (1.0 / (a.dxte[oim] + kEpsln));
      const double vhjm = vjm * a.dat[ojm] * (1.0 / (a.dytn[ojm] + kEpsln));
      const double uh = u * a.dat[os] * (1.0 / (a.dxte[os] + kEpsln));
      const double vh = v * a.dat[os] * (1.0 / (a.dytn[os] + kEpsln));
      dv = ((uh - uhim) + (vh - vhjm)) * a.datr[os];
    } else {
      const int oimjm = ojm - 1;
      const double ujm = mean_ud(a, rnts, 0, si, sj - 1);
      const double vim = mean_ud(a, rnts, 1, si - 1, sj);
      const double uimjm = mean_ud(a, rnts, 0, si - 1, sj - 1);
      const double vimjm = mean_ud(a, rnts, 1, si - 1, sj - 1);
      const double uh = u * a.dyu[os], vh = v * a.dxu[os];
      const double uhjm = ujm * a.dyu[ojm], vhjm = vjm * a.dxu[ojm];
      const double uhim = uim * a.dyu[oim], vhim = vim * a.dxu[oim];
      const double uhimjm = uimjm * a.dyu[oimjm], vhimjm = vimjm * a.dxu[oimjm];
      const double uh_bay = 0.5 * (uh + uhjm);
      const double vh_bax = 0.5 * (vh + vhim);
      const double uhim_bay = 0.5 * (uhim + uhimjm);
      const double vhjm_bax = 0.5 * (vhjm + vhimjm);
      dv = ((uh_bay - uhim_bay) + (vh_bax - vhjm_bax)) * a.datr[os];
    }
  }
  conv[o] = -dv;
}

template <bool CG, bool PR>
int run_variant(mom_gpu_ctx *ctx, const mom_gpu_bt_params_t *p, const mom_gpu_bt_variant_t *v) {
  const KDom &d = ctx->kd;
  const size_t P2 = (size_t)d.P2;
  BtvArgs<CG, PR> a{};
  a.d = d;
  a.dxu = ctx->dxu; a.dyu = ctx->dyu; a.dxur = ctx->dxur; a.dyur = ctx->dyur;
  a.dat = ctx->dat; a.datr = ctx->datr; a.dxte = ctx->dxte; a.dytn = ctx->dytn; a.f = ctx->f;
  a.kmt = ctx->kmt; a.kmu = ctx->kmu;
  a.x_tau = v->x_tau; a.xbar_tau = v->x_bar_tau; a.ud_tau = v->udrho_tau;
  a.forcing = v->forcing_bt; a.thick = v->thick; a.tmasken = v->tmasken;
  a.patm = v->patm; a.rho_g = v->rho_g; a.dpatm_dt = v->dpatm_dt;
  a.pme = v->pme; a.river = v->river; a.source = v->source; a.smooth = v->smooth;
  double *w = ctx->btv;
  a.xb = w;               // 2 planes
  a.ud = w + 2 * P2;      // 4
  a.sf = w + 6 * P2;      // 1
  a.cori1 = w + 7 * P2;   // 1
  a.cori2 = w + 8 * P2;   // 1
  a.cf = w + 9 * P2;      // 6
  a.accx = w + 15 * P2;   // 1
  a.accu = w + 16 * P2;   // 2
  a.dtbt = p->dtbt;
  a.grav = v->grav;
  a.rho0r = 1.0 / p->rho0;                                  // ocean_parameters.F90:260
  const double gamma = p->pred_corr_gamma;
  if (PR) {
    a.c_pred = p->dtbt * gamma;                             // dtbt_gamma, :1195
    a.c_corr = p->dtbt;
  } else {
    a.c_pred = p->dtbt * gamma * a.rho0r;                   // dtbt_gamma_rho0r, :1196
    a.c_corr = p->dtbt * a.rho0r;                           // dtbt_rho0r, :1197
  }
  a.predict = (PR && CG) || gamma > 0.0;                    // press_cgrid has no gamma test
  a.initsum = p->initsum_with_bar ? 1 : 0;
  hipStream_t s = ctx->stream;
  const dim3 grid((d.NI + 63) / 64, d.NJ), block(64);
  hipLaunchKernelGGL((k_btv_init<CG, PR>), grid, block, 0, s, a);
  MOM_LAUNCHED(ctx, "k_btv_init");
  const int nts = p->nts;
  for (int itime = 1; itime <= nts; ++itime) {
    const int fstau = itime % 2 + 1, fstaup1 = (itime + 1) % 2 + 1;
    const int m2 = itime % 3 + 1, m1 = (itime + 1) % 3 + 1, m0 = (itime + 2) % 3 + 1;
    hipLaunchKernelGGL((k_btv_velocity<CG, PR>), grid, block, 0, s, a, fstau, fstaup1,
                       m0, m1, m2);
    const double wgt = p->stepping_a ? 0.0 : (double)(nts - itime + 1);
    hipLaunchKernelGGL((k_btv_corrector<CG, PR>), grid, block, 0, s, a, fstau, fstaup1,
                       wgt);
  }
  MOM_LAUNCHED(ctx, "k_btv_velocity/k_btv_corrector");
  const double rntsp1 = 1.0 / ((double)nts + 1.0);          // :2804
  const double rnts = p->stepping_a ? 1.0 / (double)nts      // :2807-2811
                                    : 2.0 / ((double)nts * (double)(nts + 1));
  hipLaunchKernelGGL((k_btv_tail<CG, PR>), grid, block, 0, s, a, rntsp1, rnts,
                     v->x_bar_taup1, v->udrho_taup1, v->conv_rho_ud_t);
  MOM_LAUNCHED(ctx, "k_btv_tail");
  return MOM_GPU_OK;
}

}  // namespace

int launch_bt_variant(mom_gpu_ctx *ctx, const mom_gpu_bt_params_t *p,
                      const mom_gpu_bt_variant_t *v) {
  if (!ctx->btv) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MOM_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return mom_fail(ctx, MOM_GPU_ERR_ARG,
                      "update_ocean_barotropic_variant: the first call allocates its state; "
                      "call it once before capturing a graph");
    MOM_HIP(ctx, hipMalloc((void **)&ctx->btv, 18 * (size_t)ctx->kd.P2 * sizeof(double)));
  }
  if (v->cgrid)
    return v->pressure_based ? run_variant<true, true>(ctx, p, v) : run_variant<true, false>(ctx, p, v);
  return v->pressure_based ? run_variant<false, true>(ctx, p, v) : run_variant<false, false>(ctx, p, v);
}
