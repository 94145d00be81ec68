// Quicker tracer advection (ocean_tracer_advect.F90:1442-1586 quicker_init,
// :2538-2653 horz_advect_tracer_quicker, :2981-3031
// vert_advect_tracer_quicker) for gfx950.
#include "mom_internal.h"

namespace {

// Halo-2 semantics of dxt_quick/dyt_quick after the boundary fill and the
// halo update of quicker_init (:1494-1517): compute-domain copy, cyclic
// wrap where the axis is cyclic, copy of the nearest edge value otherwise;
// corners wrap only when both axes are cyclic.
__device__ inline double quick_metric(const KDom &d, const double *a, int gi, int gj) {
  const bool in_i = gi >= 0 && gi < d.ni, in_j = gj >= 0 && gj < d.nj;
  int si = gi, sj = gj;
  const int ci = min(max(gi, 0), d.ni - 1), cj = min(max(gj, 0), d.nj - 1);
  if (in_i && in_j) {
  } else if (!in_i && in_j) {
    si = d.cyclic_x ? wrapi(gi, d.ni) : ci;
  } else if (in_i && !in_j) {
    sj = d.cyclic_y ? wrapi(gj, d.nj) : cj;
  } else {
    if (d.cyclic_x && d.cyclic_y) { si = wrapi(gi, d.ni); sj = wrapi(gj, d.nj); }
    else { si = ci; sj = cj; }
  }
  return a[d2(d, si, sj)];
}

// Weights on (isc-1:iec, jsc-1:jec); zero elsewhere on the data domain.
__global__ void k_quicker_init(KDom d, const double *__restrict__ dxt,
                               const double *__restrict__ dyt, double *__restrict__ qw) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;  // 0..NI-1
  const int jj = blockIdx.y;                             // 0..NJ-1
  if (ii >= d.NI) return;
  const int gi = ii - 1, gj = jj - 1;
  const int o = ii + d.NI * jj;
  const int P2 = d.P2;
  double v[16];
#pragma unroll
  for (int n = 0; n < 16; ++n) v[n] = 0.0;
  if (gi >= -1 && gi <= d.ni - 1 && gj >= -1 && gj <= d.nj - 1) {
    const double xm = quick_metric(d, dxt, gi - 1, gj), x0 = quick_metric(d, dxt, gi, gj);
    const double x1 = quick_metric(d, dxt, gi + 1, gj), x2 = quick_metric(d, dxt, gi + 2, gj);
    const double ym = quick_metric(d, dyt, gi, gj - 1), y0 = quick_metric(d, dyt, gi, gj);
    const double y1 = quick_metric(d, dyt, gi, gj + 1), y2 = quick_metric(d, dyt, gi, gj + 2);
    v[0] = x1 / (x1 + x0);
    v[1] = x0 / (x1 + x0);
    v[2] = y1 / (y1 + y0);
    v[3] = y0 / (y1 + y0);
    v[4] = x0 * x1 / ((xm + 2.0 * x0 + x1) * (x0 + x1));
    v[5] = -x0 * x1 / ((x0 + x1) * (xm + x0));
    v[6] = x0 * x1 / ((xm + 2.0 * x0 + x1) * (xm + x0));
    v[7] = x0 * x1 / ((x0 + 2.0 * x1 + x2) * (x1 + x2));
    v[8] = -x0 * x1 / ((x1 + x2) * (x0 + x1));
    v[9] = x0 * x1 / ((x0 + 2.0 * x1 + x2) * (x0 + x1));
    v[10] = y0 * y1 / ((ym + 2.0 * y0 + y1) * (y0 + y1));
    v[11] = -y0 * y1 / ((y0 + y1) * (ym + y0));
    v[12] = y0 * y1 / ((ym + 2.0 * y0 + y1) * (ym + y0));
    v[13] = y0 * y1 / ((y0 + 2.0 * y1 + y2) * (y1 + y2));
    v[14] = -y0 * y1 / ((y1 + y2) * (y0 + y1));
    v[15] = y0 * y1 / ((y0 + 2.0 * y1 + y2) * (y0 + y1));
  }
#pragma unroll
  for (int n = 0; n < 16; ++n) qw[o + n * P2] = v[n];
}

__global__ void k_quicker_init_z(int nk, const double *__restrict__ dzt, double *__restrict__ qz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (k > nk) return;
  const int kp2 = min(k + 2, nk), kp1 = min(k + 1, nk), km1 = max(k - 1, 1);
  const double zm = dzt[km1 - 1], z0 = dzt[k - 1], z1 = dzt[kp1 - 1], z2 = dzt[kp2 - 1];
  qz[k - 1] = z1 / (z1 + z0);
  qz[k - 1 + nk] = z0 / (z1 + z0);
  qz[k - 1 + 2 * nk] = z0 * z1 / ((zm + 2.0 * z0 + z1) * (z0 + z1));
  qz[k - 1 + 3 * nk] = -z0 * z1 / ((z0 + z1) * (zm + z0));
  qz[k - 1 + 4 * nk] = z0 * z1 / ((zm + 2.0 * z0 + z1) * (zm + z0));
  qz[k - 1 + 5 * nk] = z0 * z1 / ((z0 + 2.0 * z1 + z2) * (z1 + z2));
  qz[k - 1 + 6 * nk] = -z0 * z1 / ((z1 + z2) * (z0 + z1));
  qz[k - 1 + 7 * nk] = z0 * z1 / ((z0 + 2.0 * z1 + z2) * (z0 + z1));
}

// horz_advect_tracer_quicker (:2538-2653), marching down the levels.  A
// work-group owns a 64-wide region of NW rows (one row per wavefront; a
// 2-point halo on each side, so 60 x (NW-4) interior points) and walks k:
// the 16 Quicker weights, dyte, dxtn and datr of a column are two-dimensional
// and stay in registers for every level, the x neighbours come by DPP lane
// shifts and the y neighbours through LDS.  Each point forms the flux through
// its east face and its north face; the cell takes the west and south ones
// from its neighbours.
//
// The inputs keep the reference's two readings: the upwind-biased stencil
// reads T(taum1) and the masks with tracer_quick / tmask_quick semantics
// (compute-domain value, cyclic wrap, zero beyond a closed axis, corners only
// when both axes are cyclic), while T(tau), the velocities, tmask_limit and
// the upwind fallback's T(taum1) and kmt are read at the face's own data-
// domain points, halo included.  Every flux and divergence is the
// reference's expression, operation for operation.
constexpr int QRX = 64, QTX = QRX - 4;

template <int NW>
__global__ __launch_bounds__(QRX *NW) void k_quicker_horz(
    KDom d, const double *__restrict__ Tm1, const double *__restrict__ Tt,
    const double *__restrict__ U, const double *__restrict__ V,
    const double *__restrict__ tlim, int limit_with_upwind,
    const int *__restrict__ kmt, const double *__restrict__ dyte_a,
    const double *__restrict__ dxtn_a, const double *__restrict__ datr_a,
    const double *__restrict__ qw, double *__restrict__ th,
    double *__restrict__ wrk1, double *__restrict__ fxo, double *__restrict__ fyo,
    int ntx, int ntiles, int per_xcd) {
  constexpr int RY = NW, TY = RY - 4;
  const int b = blockIdx.x;
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= ntiles) return;
  const int tile = strip_tile<8>(t, ntx, ntiles / ntx);
  const int tix = tile % ntx, tiy = tile / ntx;
  const int lane = threadIdx.x & (QRX - 1), row = threadIdx.x / QRX;
  const int gi = tix * QTX - 2 + lane, gj = tiy * TY - 2 + row;
  const int nk = d.nk, P2 = d.P2;

  __shared__ double s_t[RY][QRX];    // T(taum1), quick semantics
  __shared__ double s_tt[RY][QRX];   // T(tau) at the data-domain point
  __shared__ double s_fy[RY][QRX];   // flux through the north face of each point
  __shared__ int s_km[RY][QRX];

  // the point with quick semantics
  int si = gi, sj = gj;
  bool colok = true, rowok = true;
  if (si < 0 || si >= d.ni) {
    if (d.cyclic_x) si = wrapi(si, d.ni); else colok = false;
  }
  if (sj < 0 || sj >= d.nj) {
    if (d.cyclic_y) sj = wrapi(sj, d.nj); else rowok = false;
  }
  const bool ok = colok && rowok;
  const int off = ok ? d2(d, si, sj) : 0;
  const int km = ok ? kmt[off] : 0;
  s_km[row][lane] = km;
  // the point as a data-domain index (halo included), where it exists
  const bool indata = gi >= -1 && gi <= d.ni && gj >= -1 && gj <= d.nj;
  const int raw = indata ? d2(d, gi, gj) : 0;

  // faces this point forms: the east face of a compute row, the north face
  // of a compute column; the cells of the region's interior
  const bool in_j = gj >= 0 && gj < d.nj, in_i = gi >= 0 && gi < d.ni;
  const bool xface = in_j && lane >= 1 && lane <= QRX - 3 && gi >= -1 && gi <= d.ni - 1;
  const bool yface = in_i && row >= 1 && row <= RY - 3 && gj >= -1 && gj <= d.nj - 1;
  const bool cell = in_i && in_j && row >= 2 && row < RY - 2 && lane >= 2 && lane < QRX - 2;
  double qx0 = 0.0, qx1 = 0.0, cxp0 = 0.0, cxp1 = 0.0, cxp2 = 0.0, cxn0 = 0.0, cxn1 = 0.0,
         cxn2 = 0.0, dyte = 0.0;
  double qy0 = 0.0, qy1 = 0.0, cyp0 = 0.0, cyp1 = 0.0, cyp2 = 0.0, cyn0 = 0.0, cyn1 = 0.0,
         cyn2 = 0.0, dxtn = 0.0;
  if (xface) {
    const double *w = qw + raw;
    qx0 = w[0]; qx1 = w[P2];
    cxp0 = w[4 * P2]; cxp1 = w[5 * P2]; cxp2 = w[6 * P2];
    cxn0 = w[7 * P2]; cxn1 = w[8 * P2]; cxn2 = w[9 * P2];
    dyte = dyte_a[raw];
  }
  if (yface) {
    const double *w = qw + raw;
    qy0 = w[2 * P2]; qy1 = w[3 * P2];
    cyp0 = w[10 * P2]; cyp1 = w[11 * P2]; cyp2 = w[12 * P2];
    cyn0 = w[13 * P2]; cyn1 = w[14 * P2]; cyn2 = w[15 * P2];
    dxtn = dxtn_a[raw];
  }
  const double datr = cell ? datr_a[raw] : 0.0;
  const int kraw = cell ? kmt[raw] : 0;
  __syncthreads();
  // depths of the neighbours the coastal substitutions read
  const int km_w = dpp_prev(km);
  const int km_e = dpp_next(km), km_e2 = dpp_next(km_e);
  const int km_s = s_km[max(row - 1, 0)][lane];
  const int km_n = s_km[min(row + 1, RY - 1)][lane];
  const int km_n2 = s_km[min(row + 2, RY - 1)][lane];

  // A level's T(taum1), T(tau), u and v are issued during the level before
  // (unconditionally, from offset 0 where the point does not exist), so
  // that the phases of a level wait for none of them.
  double nt1 = Tm1[off], ntt = Tt[raw], nu = U[raw], nv = V[raw];
  for (int k = 1; k <= nk; ++k) {
    const int k0 = k - 1;
    const size_t pk = (size_t)P2 * k0;
    const double t1 = ok ? nt1 : 0.0;
    const double tt = indata ? ntt : 0.0;
    const double u = nu, v = nv;
    {
      const size_t pn = (size_t)P2 * min(k, nk - 1);   // the last level reloads its own
      nt1 = Tm1[off + pn];
      ntt = Tt[raw + pn];
      nu = U[raw + pn];
      nv = V[raw + pn];
    }
    s_t[row][lane] = t1;
    s_tt[row][lane] = tt;
    __syncthreads();
    const double te = lane_next(t1), te2 = lane_next(te), tw = lane_prev(t1);
    const double tte = lane_next(tt);
    const double tn = s_t[min(row + 1, RY - 1)][lane];
    const double tn2 = s_t[min(row + 2, RY - 1)][lane];
    const double ts = s_t[max(row - 1, 0)][lane];
    const double ttn = s_tt[min(row + 1, RY - 1)][lane];
    const double m0 = (k <= km) ? 1.0 : 0.0;

    // x flux through the east face of this point (:2590-2606, 2610-2622)
    double fx = 0.0;
    if (xface) {
      const double vel = dyte * u;
      const double upos = 0.5 * (vel + fabs(vel));
      const double uneg = 0.5 * (vel - fabs(vel));
      const double eastmsk = m0 * (1.0 - ((k <= km_w) ? 1.0 : 0.0));
      const double westmsk = ((k <= km_e) ? 1.0 : 0.0) * (1.0 - ((k <= km_e2) ? 1.0 : 0.0));
      fx = vel * (qx0 * tt + qx1 * tte) -
           upos * (cxp0 * te + cxp1 * t1 + cxp2 * (tw * (1.0 - eastmsk) + t1 * eastmsk)) -
           uneg * (cxn0 * (te2 * (1.0 - westmsk) + te * westmsk) + cxn1 * te + cxn2 * t1);
      if (limit_with_upwind && tlim[raw + pk] == 1.0) {
        const double up = 0.5 * (u + fabs(u)), un = 0.5 * (u - fabs(u));
        const double ma = (k <= kmt[raw]) ? 1.0 : 0.0, mb = (k <= kmt[raw + 1]) ? 1.0 : 0.0;
        fx = dyte * (up * Tm1[raw + pk] + un * Tm1[raw + 1 + pk]) * ma * mb;
      }
    }
    const double fxw = lane_prev(fx);
    // y flux through the north face (:2573-2587, 2623-2635)
    if (yface) {
      const double vel = dxtn * v;
      const double upos = 0.5 * (vel + fabs(vel));
      const double uneg = 0.5 * (vel - fabs(vel));
      const double rnormsk = m0 * (1.0 - ((k <= km_s) ? 1.0 : 0.0));
      const double soutmsk = ((k <= km_n) ? 1.0 : 0.0) * (1.0 - ((k <= km_n2) ? 1.0 : 0.0));
      double fy = vel * (qy0 * tt + qy1 * ttn) -
                  upos * (cyp0 * tn + cyp1 * t1 + cyp2 * (ts * (1.0 - rnormsk) + t1 * rnormsk)) -
                  uneg * (cyn0 * (tn2 * (1.0 - soutmsk) + tn * soutmsk) + cyn1 * tn + cyn2 * t1);
      if (limit_with_upwind && tlim[raw + pk] == 1.0) {
        const double up = 0.5 * (v + fabs(v)), un = 0.5 * (v - fabs(v));
        const double ma = (k <= kmt[raw]) ? 1.0 : 0.0, mb = (k <= kmt[raw + d.NI]) ? 1.0 : 0.0;
        fy = dxtn * (up * Tm1[raw + pk] + un * Tm1[raw + d.NI + pk]) * ma * mb;
      }
      s_fy[row][lane] = fy;
    }
    __syncthreads();

    if (cell) {
      const double fy = s_fy[row][lane], fys = s_fy[row - 1][lane];
      const double m = (k <= kraw) ? 1.0 : 0.0;
      const double res = m * (fx - fxw + fy - fys) * datr;
      const double w1 = -res;
      th[raw + pk] = th[raw + pk] + w1;
      if (wrk1) wrk1[raw + pk] = w1;
      if (fxo) {
        fxo[raw + pk] = fx;
        if (gi == 0) fxo[raw - 1 + pk] = fxw;
      }
      if (fyo) {
        fyo[raw + pk] = fy;
        if (gj == 0) fyo[raw - d.NI + pk] = fys;
      }
    }
  }
}

// One thread per column, k marching with the top flux carried (:3002-3028).
__global__ void k_quicker_vert(KDom d, const double *__restrict__ Tm1,
                               const double *__restrict__ Tt, const double *__restrict__ W,
                               const double *__restrict__ tlim, const int *__restrict__ kmt,
                               const double *__restrict__ dat, const double *__restrict__ qz,
                               double *__restrict__ th, double *__restrict__ wrk1,
                               double *__restrict__ fzo) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  const int gj = blockIdx.y;
  if (gi >= d.ni) return;
  const int nk = d.nk, P2 = d.P2;
  const int o = d2(d, gi, gj);
  const int km = kmt[o];
  const double *quick_z = qz, *czp = qz + 2 * nk, *czn = qz + 5 * nk;
  double ft1 = 0.0;
  for (int k = 1; k <= nk; ++k) {
    const int km1 = max(k - 1, 1), kp1 = min(k + 1, nk), p2 = min(k + 2, nk);
    const double vel = W[o + P2 * k];
    const double upos = 0.5 * (vel + fabs(vel));
    const double uneg = 0.5 * (vel - fabs(vel));
    const double mk0 = (k <= km) ? 1.0 : 0.0, mkp1 = (kp1 <= km) ? 1.0 : 0.0;
    const double tl = tlim ? tlim[o + P2 * (k - 1)] : 0.0;
    double ft2;
    if (tl == 1.0) {
      ft2 = (uneg * Tm1[o + P2 * (k - 1)] + upos * Tm1[o + P2 * (kp1 - 1)]) * mk0 * mkp1;
    } else {
      const double tp2 = (p2 <= km) ? 1.0 : 0.0;
      const int kp2 = (int)lrint(tp2 * (double)p2 + (1.0 - tp2) * (double)kp1);
      const double mkm1 = (km1 <= km) ? 1.0 : 0.0;
      const double upmsk = mk0 * (1.0 - mkm1);
      ft2 = vel * (quick_z[k - 1] * Tt[o + P2 * (k - 1)] + quick_z[k - 1 + nk] * Tt[o + P2 * (kp1 - 1)]) -
            uneg * (czp[k - 1] * Tm1[o + P2 * (kp1 - 1)] + czp[k - 1 + nk] * Tm1[o + P2 * (k - 1)] +
                    czp[k - 1 + 2 * nk] * Tm1[o + P2 * (km1 - 1)]) -
            upos * (czn[k - 1] * Tm1[o + P2 * (kp2 - 1)] + czn[k - 1 + nk] * Tm1[o + P2 * (kp1 - 1)] +
                    czn[k - 1 + 2 * nk] * (Tm1[o + P2 * (k - 1)] * (1.0 - upmsk) +
                                           Tm1[o + P2 * (kp1 - 1)] * upmsk));
    }
    if (fzo) fzo[o + P2 * (k - 1)] = dat[o] * ft2;
    const double res = mk0 * (ft1 - ft2);
    const double w1 = -res;
    th[o + P2 * (k - 1)] = th[o + P2 * (k - 1)] + w1;
    if (wrk1) wrk1[o + P2 * (k - 1)] = w1;
    ft1 = ft2;
  }
}

}  // namespace

int launch_quicker_init(mom_gpu_ctx *ctx) {
  const KDom &d = ctx->kd;
  hipLaunchKernelGGL(k_quicker_init, dim3((d.NI + 127) / 128, d.NJ), dim3(128), 0, ctx->stream,
                     d, ctx->dxt, ctx->dyt, ctx->qw);
  MOM_LAUNCHED(ctx, "k_quicker_init");
  hipLaunchKernelGGL(k_quicker_init_z, dim3((d.nk + 63) / 64), dim3(64), 0, ctx->stream, d.nk,
                     ctx->dzt, ctx->qzw);
  MOM_LAUNCHED(ctx, "k_quicker_init_z");
  return MOM_GPU_OK;
}

int launch_quicker_horz(mom_gpu_ctx *ctx, const double *Tm1, const double *Tt,
                        const double *u, const double *v, const double *tlim,
                        int limit_with_upwind, double *th, double *wrk1, double *fx,
                        double *fy) {
  const KDom &d = ctx->kd;
  constexpr int NW = 16;
  const int ntx = (d.ni + QTX - 1) / QTX, nty = (d.nj + NW - 4 - 1) / (NW - 4);
  const int ntiles = ntx * nty, per_xcd = (ntiles + 7) / 8;
  hipLaunchKernelGGL(k_quicker_horz<NW>, dim3(8 * per_xcd), dim3(QRX * NW), 0, ctx->stream, d,
                     Tm1, Tt, u, v, tlim, limit_with_upwind, ctx->kmt, ctx->dyte, ctx->dxtn,
                     ctx->datr, ctx->qw, th, wrk1, fx, fy, ntx, ntiles, per_xcd);
  MOM_LAUNCHED(ctx, "k_quicker_horz");
  return MOM_GPU_OK;
}

int launch_quicker_vert(mom_gpu_ctx *ctx, const double *Tm1, const double *Tt,
                        const double *w, const double *tlim, double *th, double *wrk1,
                        double *fz) {
  const KDom &d = ctx->kd;
  hipLaunchKernelGGL(k_quicker_vert, dim3((d.ni + 127) / 128, d.nj), dim3(128), 0, ctx->stream,
                     d, Tm1, Tt, w, tlim, ctx->kmt, ctx->dat, ctx->qzw, th, wrk1, fz);
  MOM_LAUNCHED(ctx, "k_quicker_vert");
  return MOM_GPU_OK;
}
