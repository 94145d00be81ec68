// Advection velocity on the B-grid: ocean_advection_velocity
// (ocean_core/ocean_advection_velocity.F90:568-672, no blobs, OBC or
// prescribed inflow), T-cell part.
//
// One thread per data-domain column marches down k.  The operators of
// ocean_core/ocean_operators.F90 (halo 1) are evaluated point-wise: BAY
// (:889-902) and BAX (:834-846) average the U-cell transport u*dy(x)u*rho_dzu
// backwards and divide by the T-face length; BDX_ET (:945-958) and BDY_NT
// (:1230-1243) difference the face transports backwards.  The west and south
// neighbours' face transports are recomputed by the same expression, so every
// value equals what the reference's array-valued functions hold, bit for bit.
#include "mom_internal.h"

namespace {

constexpr int AB = 64;

__global__ __launch_bounds__(AB) void k_adv_vel(
    KDom d, const double *__restrict__ u1, const double *__restrict__ u2,
    const double *__restrict__ rdu, const double *__restrict__ dyu,
    const double *__restrict__ dxu, const double *__restrict__ dyte,
    const double *__restrict__ dxtn, const double *__restrict__ datr,
    const int *__restrict__ kmt, const double *__restrict__ pme,
    const double *__restrict__ river, const double *__restrict__ rtend,
    const double *__restrict__ msrc, double *__restrict__ ue_out,
    double *__restrict__ vn_out, double *__restrict__ wb, double *__restrict__ div_out) {
  const int ii = blockIdx.x * AB + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  const bool hw = ii >= 1, hs = jj >= 1;
  const int ow = hw ? o - 1 : o, os = hs ? o - d.NI : o;
  const int osw = (hw && hs) ? o - d.NI - 1 : o;
  const int km = kmt[o];
  const size_t P2 = (size_t)d.P2;
  double w = -(pme[o] + river[o]);
  wb[o] = w;
  for (int k = 0; k < d.nk; ++k) {
    const size_t L = P2 * (size_t)k;
    // BAY(u1*dyu*rho_dzu)/dyte at (i,j) and (i-1,j)
    const double a_c = u1[L + o] * dyu[o] * rdu[L + o];
    const double ue_c = (hs ? 0.5 * (a_c + u1[L + os] * dyu[os] * rdu[L + os]) : 0.0) / dyte[o];
    // BAX(u2*dxu*rho_dzu)/dxtn at (i,j) and (i,j-1)
    const double b_c = u2[L + o] * dxu[o] * rdu[L + o];
    const double vn_c = (hw ? 0.5 * (b_c + u2[L + ow] * dxu[ow] * rdu[L + ow]) : 0.0) / dxtn[o];
    double bdx = 0.0, bdy = 0.0;
    if (hw) {
      const double a_w = u1[L + ow] * dyu[ow] * rdu[L + ow];
      const double ue_w =
          (hs ? 0.5 * (a_w + u1[L + osw] * dyu[osw] * rdu[L + osw]) : 0.0) / dyte[ow];
      bdx = (dyte[o] * ue_c - dyte[ow] * ue_w) * datr[o];
    }
    if (hs) {
      const double b_s = u2[L + os] * dxu[os] * rdu[L + os];
      const double vn_s =
          (hw ? 0.5 * (b_s + u2[L + osw] * dxu[osw] * rdu[L + osw]) : 0.0) / dxtn[os];
      bdy = (dxtn[o] * vn_c - dxtn[os] * vn_s) * datr[o];
    }
    const double tm = (k + 1 <= km) ? 1.0 : 0.0;
    const double dv = tm * (bdx + bdy);
    const double tmp = (rtend ? rtend[L + o] : 0.0) - (msrc ? msrc[L + o] : 0.0);
    w = (tmp + dv + w) * tm;
    ue_out[L + o] = ue_c;
    vn_out[L + o] = vn_c;
    wb[L + P2 + o] = w;
    if (div_out) div_out[L + o] = dv;
  }
}

}  // namespace

int launch_adv_vel(mom_gpu_ctx *ctx, const double *u, const double *rho_dzu, const double *pme,
                   const double *river, const double *rtend, const double *msrc, double *ue,
                   double *vn, double *wb, double *div) {
  const KDom &d = ctx->kd;
  const size_t p3 = (size_t)d.P2 * d.nk;
  hipLaunchKernelGGL(k_adv_vel, dim3((d.NI + AB - 1) / AB, d.NJ), dim3(AB), 0, ctx->stream, d, u,
                     u + p3, rho_dzu, ctx->dyu, ctx->dxu, ctx->dyte, ctx->dxtn, ctx->datr,
                     ctx->kmt, pme, river, rtend, msrc, ue, vn, wb, div);
  MOM_LAUNCHED(ctx, "k_adv_vel");
  return MOM_GPU_OK;
}
