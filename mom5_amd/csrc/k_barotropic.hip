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

#include <algorithm>
#include <cstring>

#include <algorithm>

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

// ---- temporally blocked subcycles -----------------------------------------
// One work-group owns a BRY x BRX region of the data domain and advances it
// nsub subcycles with the state (eta, U) and the accumulators in registers
// and only the exchanged quantities (U*dyu, U*dxu, surface pressure) in
// LDS.  Per subcycle the velocity's exact region shrinks by one point on
// each side (its pressure gradient reaches one point up, the divergence in
// that pressure one point down) and the height, formed from the divergence
// of the new velocity, lags one more point on the low side: after nsub
// subcycles the points from nsub + 1 above the region's low edges to nsub
// below its high edges are exact, an interior of (BRY-2*nsub-1) x
// (BRX-2*nsub-1); the rest is recomputed by the neighbouring groups.  On a
// cyclic-x domain the region addresses the
// periodic images of the compute columns: with the inputs' halo columns
// holding the wrap (what mpp_update_domains leaves), this reproduces the
// reference's evolution of every point, the halo columns of the compute
// rows excepted -- those are rewritten by the tail's halo update anyway.
constexpr int BRX = 64;          // region columns (one wavefront per row)
constexpr int BRY = 32;          // region rows
constexpr int BPT = 2;           // region rows per wavefront
constexpr int BNT = BRX * BRY / BPT;

// the halo a region keeps below its interior, and the interior's extent, for
// nsub subcycles
__host__ __device__ constexpr int bt_lo(int nsub) { return nsub + 1; }
__host__ __device__ constexpr int bt_interior(int n, int nsub) { return n - 2 * nsub - 1; }

// sf: tmask * (pme + river + source - eta_smooth) and mask: bit 0 tmask(k=1),
// bit 1 umask(k=1), formed once per call by k_bt_sfmask (constant over the
// subcycles) instead of being re-read from their six sources by every launch
struct BtIn {
  const double *rho_g, *patm, *thicku, *forcing, *f;
  const double *dxu, *dyu, *datr, *dxur, *dyur;
  const double *sf;
  const unsigned char *mask;
  int recip;   // dxur, dyur formed as 1/(dxu + epsln), 1/(dyu + epsln), not loaded
};

constexpr double EPSLN_GRID = 1.0e-40;   // shared/constants/constants.F90:123

// Per data-domain point: bit 0 where dxur or dyur is not, bit for bit, the
// reciprocal ocean_grids.F90:949-951 forms; bit 1 where a metric the blocked
// barotropic multiplies is outside LAND_BOUND in magnitude (or not a number)
// or dxu, dyu, datr carries a sign bit.  The host ORs them up.
constexpr double LAND_BOUND = 1.0e100;
__device__ inline bool bounded(double x) { return fabs(x) <= LAND_BOUND; }
__global__ void k_check_recip(KDom d, const double *__restrict__ dxu,
                              const double *__restrict__ dxur, const double *__restrict__ dyu,
                              const double *__restrict__ dyur, const double *__restrict__ datr,
                              const double *__restrict__ f, int *__restrict__ bad) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int o = ii + d.NI * blockIdx.y;
  if (ii >= d.NI) return;
  const double rx = 1.0 / (dxu[o] + EPSLN_GRID), ry = 1.0 / (dyu[o] + EPSLN_GRID);
  const bool recip_bad = (__builtin_bit_cast(unsigned long long, rx) !=
                          __builtin_bit_cast(unsigned long long, dxur[o])) ||
                         (__builtin_bit_cast(unsigned long long, ry) !=
                          __builtin_bit_cast(unsigned long long, dyur[o]));
  const bool grid_bad = !(bounded(dxu[o]) && bounded(dyu[o]) && bounded(datr[o]) &&
                          bounded(dxur[o]) && bounded(dyur[o]) && bounded(f[o]) &&
                          !signbit(dxu[o]) && !signbit(dyu[o]) && !signbit(datr[o]));
  bad[o] = (recip_bad ? 1 : 0) | (grid_bad ? 2 : 0);
}

// The per-call inputs that decide whether a point is inert under the
// subcycles (mask bit 2); null in Inert::eta: no point is.
struct Inert {
  const double *eta, *ud, *forcing, *rho_g, *patm, *thicku;
};
__global__ void k_bt_sfmask(KDom d, const double *__restrict__ pme,
                            const double *__restrict__ river, const double *__restrict__ source,
                            const double *__restrict__ eta_smooth, const int *__restrict__ kmt,
                            const int *__restrict__ kmu, double *__restrict__ sf,
                            unsigned char *__restrict__ mask, Inert x) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  const double tm = (1 <= kmt[o]) ? 1.0 : 0.0;
  const double s = tm * (pme[o] + river[o] + source[o] - eta_smooth[o]);
  sf[o] = s;
  const bool t = 1 <= kmt[o], u = 1 <= kmu[o];
  // inert: both masks 0, forcing and velocity zeros (either sign), sf a
  // zero, the fields multiplied into the pressure bounded
  const size_t P2 = (size_t)d.P2;
  const bool inert = x.eta && !t && !u && s == 0.0 && x.forcing[o] == 0.0 &&
                     x.forcing[o + P2] == 0.0 && x.ud[o] == 0.0 && x.ud[o + P2] == 0.0 &&
                     bounded(x.eta[o]) && bounded(x.rho_g[o]) && bounded(x.patm[o]) &&
                     bounded(x.thicku[o]);
  mask[o] = (unsigned char)((t ? 1 : 0) | (u ? 2 : 0) | (inert ? 4 : 0));
}

// One launch of the blocked subcycle for one domain: nsub subcycles from s0.
struct BtJob {
  KDom d;
  BtIn in;
  const double *eta_src, *ud_src;
  double *eta_dst, *ud_dst, *eb, *ub;
  int s0, nsub, nts, step_a;
  double dtbt, cp, cc;
  int xbase, xcount, ntx, ntiles;
  const int *land;   // per tile: 1 where bt_tile may take its land path (null: none)
};

__device__ __forceinline__ void bt_tile(const BtJob &J, int tile) {
  const KDom d = J.d;
  const BtIn &in = J.in;
  const double *__restrict__ eta_src = J.eta_src;
  const double *__restrict__ ud_src = J.ud_src;
  double *__restrict__ eta_dst = J.eta_dst;
  double *__restrict__ ud_dst = J.ud_dst;
  double *__restrict__ eb = J.eb;
  double *__restrict__ ub = J.ub;
  const int s0 = J.s0, nsub = J.nsub, nts = J.nts, step_a = J.step_a;
  const double dtbt = J.dtbt, cp = J.cp, cc = J.cc;
  const int xbase = J.xbase, xcount = J.xcount, ntx = J.ntx;
  __shared__ double s_ps[BRY][BRX];
  __shared__ double s_uh[BRY][BRX];
  __shared__ double s_vh[BRY][BRX];
  const int TX = bt_interior(BRX, nsub), TY = bt_interior(BRY, nsub), LO = bt_lo(nsub);
  const int tix = tile % ntx, tiy = tile / ntx;
  const int x_int0 = xbase + tix * TX, y_int0 = tiy * TY;
  const int lane = threadIdx.x % BRX, wave = threadIdx.x / BRX;
  const int P2 = d.P2;
  const bool cyc = d.cyclic_x != 0;

  if (J.land && J.land[tile]) {
    // Land path: every point of the region is inert (k_bt_sfmask bit 2).
    // There the velocity after each momentum phase is +0 (the pressure
    // gradient is x*0.0, the forcing and the old velocity zeros, dtbt > 0,
    // and the Coriolis factors finite), so U*dyu and U*dxu are +0 with
    // dxu, dyu >= +0 and the divergence is +0; the corrector adds
    // cc*(-0*0.0 + sf) = cc*sf, a signed zero, to eta, and the time sums
    // add eta and +0.  The velocity sums start the call at +0
    // (k_bt_acc_init) and every subcycle adds +0 to them in either path, so
    // they are left as they are.  The same values as the full path.
    const int vi = x_int0 - LO + lane;
    const int col = cyc ? 1 + wrapi(vi - 1, d.ni) : vi;
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      const int r = wave + (BNT / BRX) * q;
      const int j = y_int0 - LO + r;
      const bool inside = (cyc || (vi >= 0 && vi < d.NI)) && j >= 0 && j < d.NJ &&
                          lane >= LO && lane < LO + TX && vi < xbase + xcount && r >= LO &&
                          r < LO + TY;
      if (!inside) continue;
      const int o = col + d.NI * j;
      double e = eta_src[o], ebq = eb[o];
      const double s = in.sf[o];
      for (int t = 0; t < nsub; ++t) {
        e = e + cc * s;
        ebq = ebq + e;
      }
      eta_dst[o] = e;
      ud_dst[o] = 0.0;
      ud_dst[o + P2] = 0.0;
      eb[o] = ebq;
    }
    return;
  }

  int rr[BPT], off[BPT], vj[BPT];
  bool present[BPT], interior[BPT], div0[BPT], grad0[BPT];
  double eta[BPT], u1[BPT], u2[BPT], ebv[BPT], ub1[BPT], ub2[BPT];
  // tmask and umask at k = 1 as lane masks: x*mask is x where the mask is 1
  // and x*0.0 where it is 0 (the same bits as multiplying by the 0/1 double)
  bool tmb[BPT], umb[BPT];
  double sf[BPT], datr[BPT], dyu[BPT], dxu[BPT], rho_g[BPT], patm[BPT], thk[BPT];
  double F1[BPT], F2[BPT], fc[BPT], cori1[BPT], cori2[BPT], dxur[BPT], dyur[BPT];
  const int vi = x_int0 - LO + lane;
  int col = vi;
  bool colok = true;
  if (cyc) col = 1 + wrapi(vi - 1, d.ni);
  else colok = vi >= 0 && vi < d.NI;
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int r = wave + (BNT / BRX) * q;
    rr[q] = r;
    const int j = y_int0 - LO + r;
    vj[q] = j;
    present[q] = colok && j >= 0 && j < d.NJ;
    interior[q] = present[q] && lane >= LO && lane < LO + TX &&
                  vi < xbase + xcount && r >= LO && r < LO + TY;
    div0[q] = (j == 0) || (!cyc && vi == 0);
    grad0[q] = (j == d.NJ - 1) || (!cyc && vi == d.NI - 1);
    const int o = present[q] ? col + d.NI * j : 0;
    off[q] = o;
    if (present[q]) {
      const unsigned char mb = in.mask[o];
      tmb[q] = (mb & 1) != 0;
      sf[q] = in.sf[o];
      datr[q] = in.datr[o]; dyu[q] = in.dyu[o]; dxu[q] = in.dxu[o];
      rho_g[q] = in.rho_g[o]; patm[q] = in.patm[o]; thk[q] = in.thicku[o];
      F1[q] = in.forcing[o]; F2[q] = in.forcing[o + P2]; fc[q] = in.f[o];
      cori1[q] = 0.5 * dtbt * fc[q];
      cori2[q] = 1.0 / (1.0 + cori1[q] * cori1[q]);
      if (in.recip) {
        dxur[q] = 1.0 / (dxu[q] + EPSLN_GRID);
        dyur[q] = 1.0 / (dyu[q] + EPSLN_GRID);
      } else {
        dxur[q] = in.dxur[o]; dyur[q] = in.dyur[o];
      }
      umb[q] = (mb & 2) != 0;
      eta[q] = eta_src[o]; u1[q] = ud_src[o]; u2[q] = ud_src[o + P2];
      if (interior[q]) { ebv[q] = eb[o]; ub1[q] = ub[o]; ub2[q] = ub[o + P2]; }
      else { ebv[q] = ub1[q] = ub2[q] = 0.0; }
    } else {
      tmb[q] = umb[q] = false;
      sf[q] = datr[q] = dyu[q] = dxu[q] = rho_g[q] = patm[q] = thk[q] = 0.0;
      F1[q] = F2[q] = fc[q] = cori1[q] = cori2[q] = dxur[q] = dyur[q] = 0.0;
      eta[q] = u1[q] = u2[q] = ebv[q] = ub1[q] = ub2[q] = 0.0;
    }
    s_uh[r][lane] = u1[q] * dyu[q];
    s_vh[r][lane] = u2[q] * dxu[q];
  }
  __syncthreads();

  const int lm = max(lane - 1, 0), lp = min(lane + 1, BRX - 1);
  // The divergence of U*dy, U*dx at each point.  Phase A of a subcycle needs
  // the divergence of the velocity phase C of the previous subcycle formed,
  // from the same LDS values by the same operations, so it is carried over;
  // only the first subcycle of a launch forms it from the loaded velocity.
  double dv[BPT];
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int r = rr[q], rm = max(r - 1, 0);
    dv[q] = 0.0;
    if (!div0[q]) {
      const double uh_bay = 0.5 * (s_uh[r][lane] + s_uh[rm][lane]);
      const double vh_bax = 0.5 * (s_vh[r][lane] + s_vh[r][lm]);
      const double uhim_bay = 0.5 * (s_uh[r][lm] + s_uh[rm][lm]);
      const double vhjm_bax = 0.5 * (s_vh[rm][lane] + s_vh[rm][lm]);
      dv[q] = ((uh_bay - uhim_bay) + (vh_bax - vhjm_bax)) * datr[q];
    }
  }
  for (int t = 0; t < nsub; ++t) {
    const int itime = s0 + t;
    // A: predictor and surface pressure
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      const int r = rr[q];
      const double tmp = tmb[q] ? -dv[q] : -dv[q] * 0.0;
      const double es = eta[q] + cp * (tmp + sf[q]);
      s_ps[r][lane] = rho_g[q] * es + patm[q];
    }
    __syncthreads();
    // B: pressure gradient and momentum with implicit Coriolis
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      const int r = rr[q], rp = min(r + 1, BRY - 1);
      double g1 = 0.0, g2 = 0.0;
      if (!grad0[q]) {
        const double p = s_ps[r][lane], pe = s_ps[r][lp];
        const double pn = s_ps[rp][lane], pne = s_ps[rp][lp];
        const double p_fay = 0.5 * (p + pn);
        const double p_fax = 0.5 * (p + pe);
        const double pip_fay = 0.5 * (pe + pne);
        const double pjp_fax = 0.5 * (pn + pne);
        const double x1 = (pip_fay - p_fay) * dxur[q];
        const double x2 = (pjp_fax - p_fax) * dyur[q];
        g1 = umb[q] ? x1 : x1 * 0.0;
        g2 = umb[q] ? x2 : x2 * 0.0;
      }
      const double pf1 = -(thk[q] * g1);
      const double pf2 = -(thk[q] * g2);
      const double t1 = u1[q] + dtbt * (0.5 * fc[q] * u2[q] + pf1 + F1[q] + 0.0);
      const double t2 = u2[q] + dtbt * (-(0.5 * fc[q] * u1[q]) + pf2 + F2[q] + 0.0);
      u1[q] = cori2[q] * (t1 + cori1[q] * t2);
      u2[q] = cori2[q] * (t2 - cori1[q] * t1);
      // the previous phase C's reads of s_uh/s_vh finished before the barrier above
      s_uh[r][lane] = u1[q] * dyu[q];
      s_vh[r][lane] = u2[q] * dxu[q];
    }
    __syncthreads();
    // C: corrector and time-average accumulation
    const double wgt = (double)(nts - itime + 1);
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      const int r = rr[q], rm = max(r - 1, 0);
      double div = 0.0;
      if (!div0[q]) {
        const double uh_bay = 0.5 * (s_uh[r][lane] + s_uh[rm][lane]);
        const double vh_bax = 0.5 * (s_vh[r][lane] + s_vh[r][lm]);
        const double uhim_bay = 0.5 * (s_uh[r][lm] + s_uh[rm][lm]);
        const double vhjm_bax = 0.5 * (s_vh[rm][lane] + s_vh[rm][lm]);
        div = ((uh_bay - uhim_bay) + (vh_bax - vhjm_bax)) * datr[q];
      }
      dv[q] = div;
      const double tmp = tmb[q] ? -div : -div * 0.0;
      eta[q] = eta[q] + cc * (tmp + sf[q]);
      ebv[q] = ebv[q] + eta[q];
      if (step_a) {
        ub1[q] = ub1[q] + u1[q];
        ub2[q] = ub2[q] + u2[q];
      } else {
        ub1[q] = ub1[q] + wgt * u1[q];
        ub2[q] = ub2[q] + wgt * u2[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    if (interior[q]) {
      const int o = off[q];
      eta_dst[o] = eta[q];
      ud_dst[o] = u1[q];
      ud_dst[o + P2] = u2[q];
      eb[o] = ebv[q];
      ub[o] = ub1[q];
      ub[o + P2] = ub2[q];
    }
  }
}

// 1 for a tile of nsub subcycles whose whole region (the points bt_tile
// loads) is inert, 0 otherwise: one work-group a tile.
__global__ __launch_bounds__(256) void k_bt_land(KDom d, const unsigned char *__restrict__ mask,
                                                 int nsub, int xbase, int ntx, int ntiles,
                                                 int *__restrict__ land) {
  const int tile = blockIdx.x;
  if (tile >= ntiles) return;                  // the whole block
  const int TX = bt_interior(BRX, nsub), TY = bt_interior(BRY, nsub), LO = bt_lo(nsub);
  const int x_int0 = xbase + (tile % ntx) * TX, y_int0 = (tile / ntx) * TY;
  const int lane = threadIdx.x % BRX;
  const int vi = x_int0 - LO + lane;
  const bool cyc = d.cyclic_x != 0;
  const int col = cyc ? 1 + wrapi(vi - 1, d.ni) : vi;
  const bool colok = cyc || (vi >= 0 && vi < d.NI);
  int ok = 1;
  for (int r = threadIdx.x / BRX; r < BRY; r += 256 / BRX) {
    const int j = y_int0 - LO + r;
    if (colok && j >= 0 && j < d.NJ && !(mask[col + d.NI * j] & 4)) ok = 0;
  }
  ok = __syncthreads_and(ok);
  if (threadIdx.x == 0) land[tile] = ok;
}

// Tiles in XCD-aware order: blocks are dealt to the 8 XCDs round-robin, so
// each XCD takes per_xcd consecutive positions of the order of 8-wide strips
// (strip_tile) and the halo rows and columns that neighbouring tiles both
// read are fetched into one L2.
__global__ __launch_bounds__(BNT) void k_bt_block(BtJob J, int per_xcd) {
  const int b = blockIdx.x;
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= J.ntiles) return;
  bt_tile(J, strip_tile<8>(t, J.ntx, J.ntiles / J.ntx));
}

// Ensemble: launch ell of every domain that still has subcycles; the static
// per-domain part is a device table, the per-launch tile prefix an argument.
constexpr int BT_MAXDOM = 64;
struct BtDom {
  KDom d;
  BtIn in;
  const double *eta_bar_tau, *ud_tau;
  double *eta_buf[2], *ud_buf[2];
  double *eb, *ub;
  int nts, step_a;
  double dtbt, cp, cc;
  int xbase, xcount;
};
struct BtLaunch {
  int first[BT_MAXDOM + 1];   // first tile of each domain in this launch
  int ndom, ell, hsub;
};

__device__ __host__ inline void bt_shape(const BtDom &D, int ell, int hsub, int &nsub, int &ntx,
                                         int &ntiles) {
  const int left = D.nts - ell * hsub;
  nsub = left < hsub ? left : hsub;
  if (nsub <= 0) { nsub = 0; ntx = 0; ntiles = 0; return; }
  const int TX = bt_interior(BRX, nsub), TY = bt_interior(BRY, nsub);
  ntx = (D.xcount + TX - 1) / TX;
  ntiles = ntx * ((D.d.NJ + TY - 1) / TY);
}

__global__ __launch_bounds__(BNT) void k_bt_block_ens(const BtDom *__restrict__ doms, BtLaunch L) {
  const int b = blockIdx.x;
  if (b >= L.first[L.ndom]) return;
  int lo = 0, hi = L.ndom - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.first[mid] <= b) lo = mid; else hi = mid - 1;
  }
  const BtDom &D = doms[lo];
  BtJob J;
  J.d = D.d;
  J.in = D.in;
  int nsub, ntx, ntiles;
  bt_shape(D, L.ell, L.hsub, nsub, ntx, ntiles);
  J.eta_src = L.ell == 0 ? D.eta_bar_tau : D.eta_buf[(L.ell - 1) & 1];
  J.ud_src = L.ell == 0 ? D.ud_tau : D.ud_buf[(L.ell - 1) & 1];
  J.eta_dst = D.eta_buf[L.ell & 1];
  J.ud_dst = D.ud_buf[L.ell & 1];
  J.eb = D.eb;
  J.ub = D.ub;
  J.s0 = L.ell * L.hsub + 1;
  J.nsub = nsub;
  J.nts = D.nts;
  J.step_a = D.step_a;
  J.dtbt = D.dtbt;
  J.cp = D.cp;
  J.cc = D.cc;
  J.xbase = D.xbase;
  J.xcount = D.xcount;
  J.ntx = ntx;
  J.ntiles = ntiles;
  J.land = nullptr;
  bt_tile(J, b - L.first[lo]);
}

__global__ void k_bt_acc_init(KDom d, int initsum, const double *__restrict__ eta_t,
                              const double *__restrict__ eta_bar_tau, double *__restrict__ eta_bar,
                              double *__restrict__ ud) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  eta_bar[o] = initsum ? eta_bar_tau[o] : eta_t[o];
  ud[o] = 0.0;
  ud[o + d.P2] = 0.0;
}

// Ensemble tails: blockIdx.z is the domain.
struct BtTail {
  KDom d;
  int initsum;
  const double *eta_t, *eta_bar_tau;
  double *eta_bar, *ud, *conv;
  double rntsp1, rnts;
  const double *dxu, *dyu, *datr;
};

__global__ void k_bt_acc_init_ens(const BtTail *__restrict__ T) {
  const BtTail &t = T[blockIdx.z];
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= t.d.NI || jj >= t.d.NJ) return;
  const int o = ii + t.d.NI * jj;
  t.eta_bar[o] = t.initsum ? t.eta_bar_tau[o] : t.eta_t[o];
  t.ud[o] = 0.0;
  t.ud[o + t.d.P2] = 0.0;
}

__global__ void k_bt_scale_ens(const BtTail *__restrict__ T) {
  const BtTail &t = T[blockIdx.z];
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= t.d.NI || jj >= t.d.NJ) return;
  const int o = ii + t.d.NI * jj;
  t.eta_bar[o] = t.rntsp1 * t.eta_bar[o];
  t.ud[o] = t.rnts * t.ud[o];
  t.ud[o + t.d.P2] = t.rnts * t.ud[o + t.d.P2];
}

// what: 0 = eta_bar and ud (after scaling), 1 = conv
__global__ void k_bt_wrap_ens(const BtTail *__restrict__ T, int what) {
  const BtTail &t = T[blockIdx.z];
  const int jj = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (!t.d.cyclic_x || jj > t.d.NJ - 2) return;
  if (what == 1 && !t.conv) return;
  double *planes[3] = {what ? t.conv : t.eta_bar, t.ud, t.ud + t.d.P2};
  const int np = what ? 1 : 3;
  for (int c = 0; c < np; ++c) {
    double *p = planes[c] + (size_t)t.d.NI * jj;
    p[0] = p[t.d.NI - 2];
    p[t.d.NI - 1] = p[1];
  }
}

__global__ void k_bt_conv_ens(const BtTail *__restrict__ T) {
  const BtTail &t = T[blockIdx.z];
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (!t.conv || ii >= t.d.NI || jj >= t.d.NJ) return;
  t.conv[ii + t.d.NI * jj] = -div_ud(t.d, t.ud, t.ud + t.d.P2, t.dxu, t.dyu, t.datr, ii, jj);
}

// ocean_barotropic_forcing (ocean_barotropic.F90:3058-3141, default namelist,
// B-grid: tmasken = umask): the column sum of accel over the wet U-levels
// plus the surface/bottom stress, on the compute domain; a halo column or
// row takes the value of its cyclic source (what the halo update leaves), a
// closed halo is zero; then the zero friction terms are added, as there.
__global__ void k_bt_forcing(KDom d, const double *__restrict__ accel,
                             const double *__restrict__ smf, const double *__restrict__ bmf,
                             const int *__restrict__ kmu, double *__restrict__ out) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  int si = ii, sj = jj;
  bool valid = true;
  if (jj < 1 || jj > d.nj) {
    if (d.cyclic_y) sj = 1 + wrapi(jj - 1, d.nj); else valid = false;
  }
  if (ii < 1 || ii > d.ni) {
    if (d.cyclic_x) si = 1 + wrapi(ii - 1, d.ni); else valid = false;
  }
  const size_t P2 = (size_t)d.P2, P3 = P2 * (size_t)d.nk;
  const int o = ii + d.NI * jj, os = si + d.NI * sj;
  const int km = valid ? kmu[os] : 0;
  for (int n = 0; n < 2; ++n) {
    double f = 0.0;
    if (valid) {
      double wrk = 0.0;
      for (int k = 1; k <= d.nk; ++k)
        wrk = wrk + (k <= km ? 1.0 : 0.0) * accel[P3 * n + P2 * (size_t)(k - 1) + os];
      f = wrk + (1 <= km ? 1.0 : 0.0) * (smf[P2 * n + os] - bmf[P2 * n + os]);
    }
    out[P2 * n + o] = f + 0.0 + 0.0;
  }
}

struct EtaIn {
  const double *conv, *pme, *river, *source, *dpatm_dt, *ht, *x_taum1;
  const int *kmt;
  double rho0r, grav, rho0grav, dtime;
  int pressure_based;
};

// eta_and_pbot_tendency (ocean_barotropic.F90:2686-2710) at one point.
__device__ inline double eta_tendency(const EtaIn &e, int o) {
  const double tm = 1 <= e.kmt[o] ? 1.0 : 0.0;
  const double sum = e.conv[o] + e.pme[o] + e.river[o] + e.source[o];
  if (!e.pressure_based) return tm * e.rho0r * sum;
  return e.dpatm_dt[o] + e.grav * tm * sum;
}

// eta_and_pbot_tendency then eta_and_pbot_update (:2303-2363): the tendency
// on the data domain; x_taup1 = x_taum1 + dtime*tendency on the compute
// domain and, on halo points the update rewrites, at their cyclic source;
// pressure-based: pbot_t = anompb + rho0grav*ht on the data domain.
__global__ void k_eta_update(KDom d, EtaIn e, double *__restrict__ x_taup1,
                             double *__restrict__ pbot, double *__restrict__ tend) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  tend[o] = eta_tendency(e, o);
  int si, sj;
  double x;
  if (halo_source(d, ii, jj, si, sj)) {
    const int os = si + d.NI * sj;
    x = e.x_taum1[os] + e.dtime * eta_tendency(e, os);
    x_taup1[o] = x;
  } else {
    x = x_taup1[o];
  }
  if (e.pressure_based) pbot[o] = x + e.rho0grav * e.ht[o];
}

struct ClosureIn {
  const double *eta_bar, *anompb_bar, *patm, *rho_g, *ht;
  const double *dte, *dtw, *dus, *dun, *daur, *dxur, *dyur;
  const double *tmasken, *dxte, *dytn;   // tmasken != NULL: C-grid gradient
  const int *kmu;
  double rho0grav;
  int pressure_based;
};

// The field REMAP_BT_TO_BU is applied to: eta_t_bar, or rho0grav*ht +
// anompb_bar for the pressure-based closure (:2908, :2918).
__device__ inline double remap_src(const ClosureIn &c, int o) {
  return c.pressure_based ? c.rho0grav * c.ht[o] + c.anompb_bar[o] : c.eta_bar[o];
}

// umask(k=1)*REMAP_BT_TO_BU at U point (i, j) (ocean_operators.F90:468-475).
__device__ inline double remap_bu(const KDom &d, const ClosureIn &c, int i, int j) {
  if (i >= d.NI - 1 || j >= d.NJ - 1) return 0.0;
  const int o = i + d.NI * j, oe = o + 1, on = o + d.NI, one = on + 1;
  const double r = (remap_src(c, o) * c.dte[o] * c.dus[o] + remap_src(c, oe) * c.dtw[oe] * c.dus[o]
                    + remap_src(c, on) * c.dte[on] * c.dun[o]
                    + remap_src(c, one) * c.dtw[one] * c.dun[o]) * c.daur[o];
  return (1 <= c.kmu[o] ? 1.0 : 0.0) * r;
}

// ps = rho_g*eta_t_bar + patm (:2970).
__device__ inline double surf_p(const ClosureIn &c, int o) {
  return c.rho_g[o] * c.eta_bar[o] + c.patm[o];
}

// GRAD_BAROTROPIC_P, halo 1, of a pressure given by a point function:
// B-grid (ocean_operators.F90:651-660) or C-grid (:664-669).
template <class P>
__device__ inline void grad_bt(const KDom &d, const ClosureIn &c, int i, int j, P p, double &gx,
                              double &gy) {
  if (i >= d.NI - 1 || j >= d.NJ - 1) {
    gx = 0.0;
    gy = 0.0;
    return;
  }
  const int o = i + d.NI * j, oe = o + 1, on = o + d.NI, one = on + 1;
  if (c.tmasken) {
    const double pc = p(o);
    gx = (p(oe) - pc) * (1.0 / (c.dxte[o] + 1.0e-40)) * c.tmasken[o];
    gy = (p(on) - pc) * (1.0 / (c.dytn[o] + 1.0e-40)) * c.tmasken[(size_t)d.P2 + o];
    return;
  }
  const double pc = p(o), pe = p(oe), pn = p(on), pne = p(one);
  const double p_fay = 0.5 * (pc + pn);
  const double p_fax = 0.5 * (pc + pe);
  const double pip_fay = 0.5 * (pe + pne);
  const double pjp_fax = 0.5 * (pn + pne);
  const double um = 1 <= c.kmu[o] ? 1.0 : 0.0;
  gx = (pip_fay - p_fay) * c.dxur[o] * um;
  gy = (pjp_fax - p_fax) * c.dyur[o] * um;
}

// Closure of update_ocean_barotropic (ocean_barotropic.F90:2906-2921,
// 2956-2985): eta_u (or pbot_u) halo-updated, ps, grad_ps and grad_anompb.
// Neighbour values of ps are recomputed, not read, so one launch suffices.
__global__ void k_bt_closure(KDom d, ClosureIn c, double *__restrict__ eta_u,
                             double *__restrict__ ps, double *__restrict__ grad_ps,
                             double *__restrict__ grad_ab) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  int si, sj;
  halo_source(d, ii, jj, si, sj);
  eta_u[o] = remap_bu(d, c, si, sj);
  ps[o] = surf_p(c, o);
  const size_t P2 = (size_t)d.P2;
  double gx, gy;
  grad_bt(d, c, ii, jj, [&](int q) { return surf_p(c, q); }, gx, gy);
  grad_ps[o] = gx;
  grad_ps[P2 + o] = gy;
  if (grad_ab) {
    grad_bt(d, c, ii, jj, [&](int q) { return c.anompb_bar[q]; }, gx, gy);
    grad_ab[o] = gx;
    grad_ab[P2 + o] = gy;
  }
}

}  // namespace

// Sets ctx->bt_recip: whether the context's dxur and dyur are the reciprocals
// the blocked barotropic may form itself (then it loads two planes fewer).
int launch_check_recip(mom_gpu_ctx *ctx) {
  const KDom &d = ctx->kd;
  const size_t n = (size_t)d.P2;
  ctx->bt_recip = 0;
  int *dev = nullptr;
  MOM_HIP(ctx, hipMalloc(&dev, n * sizeof(int)));
  hipLaunchKernelGGL(k_check_recip, dim3((d.NI + 127) / 128, d.NJ), dim3(128), 0, ctx->stream, d,
                     ctx->dxu, ctx->dxur, ctx->dyu, ctx->dyur, ctx->datr, ctx->f, dev);
  hipError_t le = hipGetLastError();
  int *host = (int *)malloc(n * sizeof(int));
  hipError_t e1 = host ? hipMemcpyAsync(host, dev, n * sizeof(int), hipMemcpyDeviceToHost,
                                        ctx->stream)
                       : hipErrorOutOfMemory;
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dev);
  if (le != hipSuccess || e1 != hipSuccess || e2 != hipSuccess) {
    free(host);
    return mom_fail(ctx, MOM_GPU_ERR_HIP, "reciprocal check failed: %s",
                    hipGetErrorString(le != hipSuccess ? le : (e1 != hipSuccess ? e1 : e2)));
  }
  int bad = 0;
  for (size_t q = 0; q < n; ++q) bad |= host[q];
  free(host);
  ctx->bt_recip = (bad & 1) == 0;
  ctx->bt_grid_ok = (bad & 2) == 0;
  return MOM_GPU_OK;
}

int launch_barotropic_blocked(mom_gpu_ctx *ctx, const mom_gpu_bt_params_t *p,
                              const double *eta_t, const double *eta_bar_tau,
                              const double *ud_tau, const double *forcing, const double *thicku,
                              const double *patm, const double *pme, const double *river,
                              const double *source, const double *eta_smooth,
                              const double *rho_g, double *eta_bar, double *ud, double *conv,
                              int hsub) {
  const KDom &d = ctx->kd;
  const size_t P2 = (size_t)d.P2;
  hipStream_t s = ctx->stream;
  const double rho0r = 1.0 / p->rho0;
  const double c_pred = p->dtbt * p->pred_corr_gamma * rho0r;   // dtbt_gamma_rho0r, :1196
  const double c_corr = p->dtbt * rho0r;                        // dtbt_rho0r, :1197
  const dim3 grid((d.NI + 255) / 256, d.NJ), block(256);
  hipLaunchKernelGGL(k_bt_acc_init, grid, block, 0, s, d, p->initsum_with_bar, eta_t,
                     eta_bar_tau, eta_bar, ud);
  MOM_LAUNCHED(ctx, "k_bt_acc_init");
  // sf and the masks in planes 6 and 7 of the subcycle buffer (0-5 hold the
  // ping-pong state)
  double *sfp = ctx->bt + 6 * P2;
  unsigned char *maskp = (unsigned char *)(ctx->bt + 7 * P2);
  // the land path's conditions that do not depend on the data
  const bool land_ok = ctx->tune.bt_land_tiles && ctx->bt_grid_ok && p->dtbt > 0.0 &&
                       std::isfinite(c_pred) && std::isfinite(c_corr);
  const Inert inert = land_ok ? Inert{eta_bar_tau, ud_tau, forcing, rho_g, patm, thicku} : Inert{};
  hipLaunchKernelGGL(k_bt_sfmask, grid, block, 0, s, d, pme, river, source, eta_smooth,
                     ctx->kmt, ctx->kmu, sfp, maskp, inert);
  BtIn in{rho_g, patm, thicku, forcing, ctx->f, ctx->dxu, ctx->dyu, ctx->datr, ctx->dxur,
          ctx->dyur, sfp, maskp, ctx->bt_recip};
  // tile flags in plane 8 (unused by the blocked path): at most one int a
  // point, as a tile's interior holds at least one point
  int *landp = land_ok ? (int *)(ctx->bt + 8 * P2) : nullptr;
  int land_nsub = 0;
  const int xbase = d.cyclic_x ? 1 : 0, xcount = d.cyclic_x ? d.ni : d.NI;
  double *eta_buf[2] = {ctx->bt, ctx->bt + P2};
  double *ud_buf[2] = {ctx->bt + 2 * P2, ctx->bt + 4 * P2};
  const double *eta_src = eta_bar_tau, *ud_src = ud_tau;
  int which = 0;
  for (int s0 = 1; s0 <= p->nts; s0 += hsub) {
    const int nsub = std::min(hsub, p->nts - s0 + 1);
    const int TX = bt_interior(BRX, nsub), TY = bt_interior(BRY, nsub);
    const int ntx = (xcount + TX - 1) / TX, nty = (d.NJ + TY - 1) / TY;
    if (landp && nsub != land_nsub) {
      // stream order: the launches that read the previous flags are done
      hipLaunchKernelGGL(k_bt_land, dim3(ntx * nty), dim3(256), 0, s, d, maskp, nsub, xbase,
                         ntx, ntx * nty, landp);
      MOM_LAUNCHED(ctx, "k_bt_land");
      land_nsub = nsub;
    }
    const BtJob J{d, in, eta_src, ud_src, eta_buf[which], ud_buf[which], eta_bar, ud, s0, nsub,
                  p->nts, p->stepping_a, p->dtbt, c_pred, c_corr, xbase, xcount, ntx, ntx * nty,
                  landp};
    const int per_xcd = (ntx * nty + 7) / 8;
    hipLaunchKernelGGL(k_bt_block, dim3(8 * per_xcd), dim3(BNT), 0, s, J, per_xcd);
    MOM_LAUNCHED(ctx, "k_bt_block");
    eta_src = eta_buf[which];
    ud_src = ud_buf[which];
    which ^= 1;
  }
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

int launch_barotropic(mom_gpu_ctx *ctx, const mom_gpu_bt_params_t *p, const double *eta_t,
                      const double *eta_bar_tau, const double *ud_tau, const double *forcing,
                      const double *thicku, const double *patm, const double *pme,
                      const double *river, const double *source, const double *eta_smooth,
                      const double *rho_g, double *eta_bar, double *ud, double *conv) {
  if (ctx->tune.bt_subcycles_per_launch > 0)
    return launch_barotropic_blocked(ctx, p, eta_t, eta_bar_tau, ud_tau, forcing, thicku, patm,
                                     pme, river, source, eta_smooth, rho_g, eta_bar, ud, conv,
                                     ctx->tune.bt_subcycles_per_launch);
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

// Barotropic subcycles of every domain of an ensemble: launch ell advances
// hsub subcycles of each domain that has any left (its own nts, dtbt).
int launch_barotropic_ensemble(mom_gpu_ensemble *e, const mom_gpu_bt_params_t *p,
                               const mom_gpu_bt_fields_t *f) {
  const int nd = (int)e->ctx.size();
  if (nd > BT_MAXDOM) return mom_fail(e->lead, MOM_GPU_ERR_ARG, "ensemble: too many domains");
  const int hsub = e->lead->tune.bt_subcycles_per_launch > 0 ? e->lead->tune.bt_subcycles_per_launch
                                                             : 4;
  std::vector<BtDom> doms(nd);
  std::vector<BtTail> tails(nd);
  std::memset((void *)doms.data(), 0, sizeof(BtDom) * nd);
  std::memset((void *)tails.data(), 0, sizeof(BtTail) * nd);
  int maxNI = 1, maxNJ = 1, nlaunch = 0;
  for (int k = 0; k < nd; ++k) {
    mom_gpu_ctx *ctx = e->ctx[k];
    const KDom &d = ctx->kd;
    const size_t P2 = (size_t)d.P2;
    const mom_gpu_bt_fields_t &x = f[k];
    const double rho0r = 1.0 / p[k].rho0;
    BtDom &D = doms[k];
    D.d = d;
    double *sfp = ctx->bt + 6 * P2;
    unsigned char *maskp = (unsigned char *)(ctx->bt + 7 * P2);
    hipLaunchKernelGGL(k_bt_sfmask, dim3((d.NI + 255) / 256, d.NJ), dim3(256), 0,
                       e->lead->stream, d, x.pme, x.river, x.source, x.eta_smooth, ctx->kmt,
                       ctx->kmu, sfp, maskp, Inert{});
    D.in = BtIn{x.rho_g, x.patm, x.thicku, x.forcing_bt, ctx->f, ctx->dxu, ctx->dyu, ctx->datr,
                ctx->dxur, ctx->dyur, sfp, maskp, ctx->bt_recip};
    D.eta_bar_tau = x.eta_t_bar_tau;
    D.ud_tau = x.udrho_tau;
    D.eta_buf[0] = ctx->bt;
    D.eta_buf[1] = ctx->bt + P2;
    D.ud_buf[0] = ctx->bt + 2 * P2;
    D.ud_buf[1] = ctx->bt + 4 * P2;
    D.eb = x.eta_t_bar_taup1;
    D.ub = x.udrho_taup1;
    D.nts = p[k].nts;
    D.step_a = p[k].stepping_a;
    D.dtbt = p[k].dtbt;
    D.cp = p[k].dtbt * p[k].pred_corr_gamma * rho0r;   // dtbt_gamma_rho0r, :1196
    D.cc = p[k].dtbt * rho0r;                          // dtbt_rho0r, :1197
    D.xbase = d.cyclic_x ? 1 : 0;
    D.xcount = d.cyclic_x ? d.ni : d.NI;
    BtTail &T = tails[k];
    T.d = d;
    T.initsum = p[k].initsum_with_bar;
    T.eta_t = x.eta_t_tau;
    T.eta_bar_tau = x.eta_t_bar_tau;
    T.eta_bar = x.eta_t_bar_taup1;
    T.ud = x.udrho_taup1;
    T.conv = x.conv_rho_ud_t;
    const double nts = (double)p[k].nts;
    T.rntsp1 = 1.0 / (nts + 1.0);
    T.rnts = p[k].stepping_a ? 1.0 / nts : 2.0 / (nts * (double)(p[k].nts + 1));
    T.dxu = ctx->dxu;
    T.dyu = ctx->dyu;
    T.datr = ctx->datr;
    maxNI = std::max(maxNI, d.NI);
    maxNJ = std::max(maxNJ, d.NJ);
    nlaunch = std::max(nlaunch, (p[k].nts + hsub - 1) / hsub);
  }
  const BtDom *ddev = (const BtDom *)ens_table(e, doms.data(), sizeof(BtDom) * nd);
  if (!ddev) return MOM_GPU_ERR_HIP;
  const BtTail *tdev = (const BtTail *)ens_table(e, tails.data(), sizeof(BtTail) * nd);
  if (!tdev) return MOM_GPU_ERR_HIP;
  hipStream_t s = e->lead->stream;
  const dim3 g2((maxNI + 255) / 256, maxNJ, nd), b2(256);
  hipLaunchKernelGGL(k_bt_acc_init_ens, g2, b2, 0, s, tdev);
  MOM_LAUNCHED(e->lead, "k_bt_acc_init_ens");
  for (int ell = 0; ell < nlaunch; ++ell) {
    BtLaunch L{};
    L.ndom = nd;
    L.ell = ell;
    L.hsub = hsub;
    int total = 0;
    for (int k = 0; k < nd; ++k) {
      int nsub, ntx, ntiles;
      bt_shape(doms[k], ell, hsub, nsub, ntx, ntiles);
      L.first[k] = total;
      total += ntiles;
    }
    L.first[nd] = total;
    if (total == 0) continue;
    hipLaunchKernelGGL(k_bt_block_ens, dim3(total), dim3(BNT), 0, s, ddev, L);
    MOM_LAUNCHED(e->lead, "k_bt_block_ens");
  }
  hipLaunchKernelGGL(k_bt_scale_ens, g2, b2, 0, s, tdev);
  const dim3 gw((maxNJ + 255) / 256, 1, nd);
  hipLaunchKernelGGL(k_bt_wrap_ens, gw, b2, 0, s, tdev, 0);
  hipLaunchKernelGGL(k_bt_conv_ens, g2, b2, 0, s, tdev);
  hipLaunchKernelGGL(k_bt_wrap_ens, gw, b2, 0, s, tdev, 1);
  MOM_LAUNCHED(e->lead, "k_bt_tail_ens");
  return MOM_GPU_OK;
}

int launch_bt_forcing(mom_gpu_ctx *ctx, const double *accel, const double *smf, const double *bmf,
                      double *forcing_bt) {
  const KDom &d = ctx->kd;
  hipLaunchKernelGGL(k_bt_forcing, dim3((d.NI + 63) / 64, d.NJ), dim3(64), 0, ctx->stream, d,
                     accel, smf, bmf, ctx->kmu, forcing_bt);
  MOM_LAUNCHED(ctx, "k_bt_forcing");
  return MOM_GPU_OK;
}

int launch_eta_update(mom_gpu_ctx *ctx, int pressure_based, double rho0, double grav,
                      double dtime, const double *conv, const double *pme, const double *river,
                      const double *source, const double *dpatm_dt, const double *ht,
                      const double *x_taum1, double *x_taup1, double *pbot, double *tend) {
  const KDom &d = ctx->kd;
  EtaIn e{conv, pme, river, source, dpatm_dt, ht, x_taum1, ctx->kmt,
          1.0 / rho0, grav, rho0 * grav, dtime, pressure_based};
  hipLaunchKernelGGL(k_eta_update, dim3((d.NI + 63) / 64, d.NJ), dim3(64), 0, ctx->stream, d, e,
                     x_taup1, pbot, tend);
  MOM_LAUNCHED(ctx, "k_eta_update");
  return MOM_GPU_OK;
}

int launch_bt_closure(mom_gpu_ctx *ctx, int pressure_based, double rho0, double grav,
                      const double *eta_bar, const double *anompb_bar, const double *patm,
                      const double *rho_g, const double *ht, const double *tmasken,
                      double *eta_u, double *ps, double *grad_ps, double *grad_ab) {
  const KDom &d = ctx->kd;
  const double *rm = ctx->remap;
  const size_t P2 = (size_t)d.P2;
  ClosureIn c{eta_bar, anompb_bar, patm, rho_g, ht, rm, rm + P2, rm + 2 * P2, rm + 3 * P2,
              rm + 4 * P2, ctx->dxur, ctx->dyur, tmasken, ctx->dxte, ctx->dytn, ctx->kmu,
              rho0 * grav, pressure_based};
  hipLaunchKernelGGL(k_bt_closure, dim3((d.NI + 63) / 64, d.NJ), dim3(64), 0, ctx->stream, d, c,
                     eta_u, ps, grad_ps, grad_ab);
  MOM_LAUNCHED(ctx, "k_bt_closure");
  return MOM_GPU_OK;
}
