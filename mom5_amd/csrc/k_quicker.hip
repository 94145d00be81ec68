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

// tracer_quick / tmask_quick semantics: compute-domain value, cyclic wrap,
// zero outside a closed axis; corners only when both axes are cyclic.
__device__ inline bool h2src(const KDom &d, int gi, int gj, int &o) {
  const bool in_i = gi >= 0 && gi < d.ni, in_j = gj >= 0 && gj < d.nj;
  if (!in_i && !d.cyclic_x) return false;
  if (!in_j && !d.cyclic_y) return false;
  if (!in_i && !in_j && !(d.cyclic_x && d.cyclic_y)) return false;
  o = d2(d, in_i ? gi : wrapi(gi, d.ni), in_j ? gj : wrapi(gj, d.nj));
  return true;
}

struct H2Val {
  double t, m;
};

__device__ inline H2Val h2val(const KDom &d, const double *Tm1, const int *kmt,
                              int gi, int gj, int k) {
  int o;
  if (!h2src(d, gi, gj, o)) return {0.0, 0.0};
  return {Tm1[o + d.P2 * (k - 1)], (k <= kmt[o]) ? 1.0 : 0.0};
}

constexpr int QBX = 64, QBY = 4;

// One k level of one 64x4 tile per block; fluxes of the tile's faces in LDS.
__global__ __launch_bounds__(QBX *QBY) void k_quicker_horz(
    KDom d, const double *__restrict__ Tm1, const double *__restrict__ Tt,
    const double *__restrict__ U, const double *__restrict__ V,
    const double *__restrict__ tlim, int limit_with_upwind,
    const int *__restrict__ kmt, const double *__restrict__ dyte,
    const double *__restrict__ dxtn, const double *__restrict__ datr,
    const double *__restrict__ qw, double *__restrict__ th,
    double *__restrict__ wrk1, double *__restrict__ fxo, double *__restrict__ fyo) {
  __shared__ double s_fx[QBY][QBX + 1];
  __shared__ double s_fy[QBY + 1][QBX];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int gi = blockIdx.x * QBX + tx, gj = blockIdx.y * QBY + ty;
  const int k = blockIdx.z + 1;
  const int P2 = d.P2, k0 = k - 1;
  const double *quick_x = qw, *quick_y = qw + 2 * P2, *cxp = qw + 4 * P2;
  const double *cxn = qw + 7 * P2, *cyp = qw + 10 * P2, *cyn = qw + 13 * P2;

  // x flux through the face east of column fi, row gj (reference :2590-2606)
  auto xflux = [&](int fi) -> double {
    const int o = d2(d, fi, gj);
    const double vel = dyte[o] * U[o + P2 * k0];
    const double upos = 0.5 * (vel + fabs(vel));
    const double uneg = 0.5 * (vel - fabs(vel));
    const H2Val a_m1 = h2val(d, Tm1, kmt, fi - 1, gj, k), a_0 = h2val(d, Tm1, kmt, fi, gj, k);
    const H2Val a_p1 = h2val(d, Tm1, kmt, fi + 1, gj, k), a_p2 = h2val(d, Tm1, kmt, fi + 2, gj, k);
    const double eastmsk = a_0.m * (1.0 - a_m1.m);
    const double westmsk = a_p1.m * (1.0 - a_p2.m);
    double fx = vel * (quick_x[o] * Tt[o + P2 * k0] + quick_x[o + P2] * Tt[o + 1 + P2 * k0]) -
                upos * (cxp[o] * a_p1.t + cxp[o + P2] * a_0.t +
                        cxp[o + 2 * P2] * (a_m1.t * (1.0 - eastmsk) + a_0.t * eastmsk)) -
                uneg * (cxn[o] * (a_p2.t * (1.0 - westmsk) + a_p1.t * westmsk) +
                        cxn[o + P2] * a_p1.t + cxn[o + 2 * P2] * a_0.t);
    if (limit_with_upwind && tlim[o + P2 * k0] == 1.0) {
      const double u = U[o + P2 * k0];
      const double up = 0.5 * (u + fabs(u)), un = 0.5 * (u - fabs(u));
      const double m0 = (k <= kmt[o]) ? 1.0 : 0.0, m1 = (k <= kmt[o + 1]) ? 1.0 : 0.0;
      fx = dyte[o] * (up * Tm1[o + P2 * k0] + un * Tm1[o + 1 + P2 * k0]) * m0 * m1;
    }
    return fx;
  };
  // y flux through the face north of row fj, column gi (:2573-2587)
  auto yflux = [&](int fj) -> double {
    const int o = d2(d, gi, fj);
    const double vel = dxtn[o] * V[o + P2 * k0];
    const double upos = 0.5 * (vel + fabs(vel));
    const double uneg = 0.5 * (vel - fabs(vel));
    const H2Val a_m1 = h2val(d, Tm1, kmt, gi, fj - 1, k), a_0 = h2val(d, Tm1, kmt, gi, fj, k);
    const H2Val a_p1 = h2val(d, Tm1, kmt, gi, fj + 1, k), a_p2 = h2val(d, Tm1, kmt, gi, fj + 2, k);
    const double rnormsk = a_0.m * (1.0 - a_m1.m);
    const double soutmsk = a_p1.m * (1.0 - a_p2.m);
    double fy = vel * (quick_y[o] * Tt[o + P2 * k0] + quick_y[o + P2] * Tt[o + d.NI + P2 * k0]) -
                upos * (cyp[o] * a_p1.t + cyp[o + P2] * a_0.t +
                        cyp[o + 2 * P2] * (a_m1.t * (1.0 - rnormsk) + a_0.t * rnormsk)) -
                uneg * (cyn[o] * (a_p2.t * (1.0 - soutmsk) + a_p1.t * soutmsk) +
                        cyn[o + P2] * a_p1.t + cyn[o + 2 * P2] * a_0.t);
    if (limit_with_upwind && tlim[o + P2 * k0] == 1.0) {
      const double v = V[o + P2 * k0];
      const double up = 0.5 * (v + fabs(v)), un = 0.5 * (v - fabs(v));
      const double m0 = (k <= kmt[o]) ? 1.0 : 0.0, m1 = (k <= kmt[o + d.NI]) ? 1.0 : 0.0;
      fy = dxtn[o] * (up * Tm1[o + P2 * k0] + un * Tm1[o + d.NI + P2 * k0]) * m0 * m1;
    }
    return fy;
  };

  const bool in = gi < d.ni && gj < d.nj;
  if (in) {
    s_fx[ty][tx + 1] = xflux(gi);
    s_fy[ty + 1][tx] = yflux(gj);
    if (tx == 0) s_fx[ty][0] = xflux(gi - 1);
    if (ty == 0) s_fy[0][tx] = yflux(gj - 1);
  }
  __syncthreads();
  if (!in) return;
  const int o = d2(d, gi, gj);
  const double fx = s_fx[ty][tx + 1], fxw = s_fx[ty][tx];
  const double fy = s_fy[ty + 1][tx], fys = s_fy[ty][tx];
  const double m = (k <= kmt[o]) ? 1.0 : 0.0;
  const double res = m * (fx - fxw + fy - fys) * datr[o];
  const double w1 = -res;
  th[o + P2 * k0] = th[o + P2 * k0] + w1;
  if (wrk1) wrk1[o + P2 * k0] = w1;
  if (fxo) {
    fxo[o + P2 * k0] = fx;
    if (gi == 0) fxo[o - 1 + P2 * k0] = fxw;
  }
  if (fyo) {
    fyo[o + P2 * k0] = fy;
    if (gj == 0) fyo[o - d.NI + P2 * k0] = fys;
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
  const dim3 grid((d.ni + QBX - 1) / QBX, (d.nj + QBY - 1) / QBY, d.nk);
  hipLaunchKernelGGL(k_quicker_horz, grid, dim3(QBX, QBY), 0, ctx->stream, d, Tm1, Tt, u, v,
                     tlim, limit_with_upwind, ctx->kmt, ctx->dyte, ctx->dxtn, ctx->datr, ctx->qw,
                     th, wrk1, fx, fy);
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
