// Batched step for Quicker tracers: horz_advect_tracer_quicker
// (ocean_tracer_advect.F90:2538-2653) and vert_advect_tracer_quicker
// (:2981-3031) into th_tendency, with the tracer's tmask_limit
// (ocean_tracer.F90:2788-2892) formed in the kernel from T(tau), then the
// explicit update (ocean_tracer.F90:2341-2350), in one k-marching pass.
//
// Two-level time stepping: T(taum1) = T(tau), one input array.  Every
// value outside the compute domain is read with MOM's halo semantics after
// the halo updates the reference makes (cyclic wrap, zero beyond a closed
// axis), as the batched MDFL pass does.  The arithmetic of every face flux,
// divergence, limit flag and update is the reference's, operation for
// operation, so T(taup1) equals bit for bit the per-tracer sequence
// (compute_tmask_limit, horz_advect_tracer and vert_advect_tracer into a
// zeroed th_tendency, the explicit update) on halo-updated inputs.
//
// Work-group layout as the MDFL pass: 64-wide region rows, one row per
// wavefront, a 2-cell halo (the Quicker stencil reaches i-1..i+2 for the
// face east of i), x neighbours through DPP lane shifts, y neighbours
// through LDS.  The 16 Quicker weights of a column are two-dimensional and
// stay in registers for all levels.
#include "mom_internal.h"

#include <vector>

namespace {

constexpr int RX = 64;
constexpr int TX = RX - 4;

__device__ inline int dpp_prev(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xf, 0xf, true); }
__device__ inline int dpp_next(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xf, 0xf, true); }
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

__device__ inline double mk(int k, int kmt) { return (k <= kmt) ? 1.0 : 0.0; }
__device__ inline bool outside(double x, double lo, double hi) { return x < lo || x > hi; }

struct QStepJob {
  KDom d;
  const double *T;      // T(taum1) = T(tau)
  const double *Tp;     // tmask_limit_ts_same partner's T(tau), or null
  double *out;          // T(taup1)
  const double *th;     // other th_tendency terms, or null
  const double *U, *V, *W, *R, *Rm1, *Rr;
  const int *kmt;
  const double *dyte, *dxtn, *datr, *qw, *qz;
  double dtime, lo, hi, plo, phi;
  int lwu;              // limit_with_upwind
  int ntx, ntiles;
};

template <int NW, bool PART>
__global__ __launch_bounds__(RX *NW) void k_quick_step(QStepJob J, int per_xcd) {
  const int b = blockIdx.x;
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= J.ntiles) return;
  const KDom d = J.d;
  constexpr int RY = NW, TY = RY - 4;
  const int tile = strip_tile<8>(t, J.ntx, J.ntiles / J.ntx);
  const int tix = tile % J.ntx, tiy = tile / J.ntx;
  const int lane = threadIdx.x & (RX - 1), row = threadIdx.x / RX;
  const int gi = tix * TX - 2 + lane, gj = tiy * TY - 2 + row;
  const int nk = d.nk, P2 = d.P2;
  const double *__restrict__ T = J.T;
  const double *__restrict__ Tp = J.Tp;
  const double dtime = J.dtime;

  __shared__ double s_t[RY][RX];     // T(k) of the region, for y neighbours
  __shared__ double s_tp[PART ? RY : 1][RX];
  __shared__ double s_fy[RY][RX];    // flux through the north face of each point
  __shared__ int s_km[RY][RX];

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
  const int km = ok ? J.kmt[off] : 0;
  s_km[row][lane] = km;

  // faces this point computes: east face of the point (x), north face (y)
  const bool xface = rowok && lane >= 1 && lane <= RX - 3 && gi >= -1 && gi <= d.ni - 1;
  const bool yface = row >= 1 && row <= RY - 3 && lane >= 2 && lane < RX - 2 && gi < d.ni &&
                     gj >= -1 && gj <= d.nj - 1;
  const bool cell = row >= 2 && row < RY - 2 && lane >= 2 && lane < RX - 2 && gi < d.ni &&
                    gj < d.nj;
  const int ox = d2(d, gi, sj), oy = d2(d, gi, gj);
  // Quicker weights (quicker_init), data-domain planes
  double qx0 = 0.0, qx1 = 0.0, cxp0 = 0.0, cxp1 = 0.0, cxp2 = 0.0, cxn0 = 0.0, cxn1 = 0.0,
         cxn2 = 0.0, dyte = 0.0;
  double qy0 = 0.0, qy1 = 0.0, cyp0 = 0.0, cyp1 = 0.0, cyp2 = 0.0, cyn0 = 0.0, cyn1 = 0.0,
         cyn2 = 0.0, dxtn = 0.0;
  if (xface) {
    const double *w = J.qw + ox;
    qx0 = w[0]; qx1 = w[P2];
    cxp0 = w[4 * P2]; cxp1 = w[5 * P2]; cxp2 = w[6 * P2];
    cxn0 = w[7 * P2]; cxn1 = w[8 * P2]; cxn2 = w[9 * P2];
    dyte = J.dyte[ox];
  }
  if (yface) {
    const double *w = J.qw + oy;
    qy0 = w[2 * P2]; qy1 = w[3 * P2];
    cyp0 = w[10 * P2]; cyp1 = w[11 * P2]; cyp2 = w[12 * P2];
    cyn0 = w[13 * P2]; cyn1 = w[14 * P2]; cyn2 = w[15 * P2];
    dxtn = J.dxtn[oy];
  }
  const double datr = cell ? J.datr[off] : 0.0;
  __syncthreads();
  // depths of the neighbours the coastal substitutions and limits read
  const int km_w = dpp_prev(km);
  const int km_e = dpp_next(km), km_e2 = dpp_next(km_e);
  const int km_s = s_km[max(row - 1, 0)][lane];
  const int km_n = s_km[min(row + 1, RY - 1)][lane];
  const int km_n2 = s_km[min(row + 2, RY - 1)][lane];

  // the k window of T: km1, k, kp1, min(k+2, nk) (clamped as the reference)
  double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0, tp1 = 0.0, tp2 = 0.0;
  if (ok) {
    t1 = T[off];
    t0 = t1;
    t2 = T[off + P2 * (min(2, nk) - 1)];
    t3 = T[off + P2 * (min(3, nk) - 1)];
    if constexpr (PART) {
      tp1 = Tp[off];
      tp2 = Tp[off + P2 * (min(2, nk) - 1)];
    }
  }
  const double *quick_z = J.qz, *czp = J.qz + 2 * nk, *czn = J.qz + 5 * nk;
  double ft1 = 0.0;

  for (int k = 1; k <= nk; ++k) {
    const int k0 = k - 1;
    const int km1 = max(k - 1, 1), kp1 = min(k + 1, nk), p2 = min(k + 2, nk);
    s_t[row][lane] = t1;
    if constexpr (PART) s_tp[row][lane] = tp1;
    __syncthreads();
    const double te = lane_next(t1), te2 = lane_next(te), tw = lane_prev(t1);
    const double tn = s_t[min(row + 1, RY - 1)][lane];
    const double tn2 = s_t[min(row + 2, RY - 1)][lane];
    const double ts = s_t[max(row - 1, 0)][lane];
    const double m0 = mk(k, km);

    // tmask_limit(i,j,k) of this point (ocean_tracer.F90:2803-2850, 2858-2869)
    double L = 0.0;
    {
      const double me = mk(k, km_e), mn = mk(k, km_n), mb = (k < nk) ? mk(k + 1, km) : 0.0;
      if (m0 == 1.0 && outside(t1, J.lo, J.hi)) L = m0;
      if (me == 1.0 && outside(te, J.lo, J.hi)) L = m0;
      if (mn == 1.0 && outside(tn, J.lo, J.hi)) L = m0;
      if (k < nk && mb == 1.0 && outside(t2, J.lo, J.hi)) L = m0;
      if constexpr (PART) {
        double Lp = 0.0;
        const double pe = lane_next(tp1), pn = s_tp[min(row + 1, RY - 1)][lane];
        if (m0 == 1.0 && outside(tp1, J.plo, J.phi)) Lp = m0;
        if (me == 1.0 && outside(pe, J.plo, J.phi)) Lp = m0;
        if (mn == 1.0 && outside(pn, J.plo, J.phi)) Lp = m0;
        if (k < nk && mb == 1.0 && outside(tp2, J.plo, J.phi)) Lp = m0;
        if (L == 1.0 || Lp == 1.0) L = 1.0;
      }
    }

    // x flux through the east face of this point (:2590-2606, 2610-2622)
    double fx = 0.0;
    if (xface) {
      const double u = J.U[ox + P2 * k0];
      const double vel = dyte * u;
      const double upos = 0.5 * (vel + fabs(vel));
      const double uneg = 0.5 * (vel - fabs(vel));
      const double eastmsk = m0 * (1.0 - mk(k, km_w));
      const double westmsk = mk(k, km_e) * (1.0 - mk(k, km_e2));
      fx = vel * (qx0 * t1 + qx1 * te) -
           upos * (cxp0 * te + cxp1 * t1 + cxp2 * (tw * (1.0 - eastmsk) + t1 * eastmsk)) -
           uneg * (cxn0 * (te2 * (1.0 - westmsk) + te * westmsk) + cxn1 * te + cxn2 * t1);
      if (J.lwu && L == 1.0) {
        const double up = 0.5 * (u + fabs(u)), un = 0.5 * (u - fabs(u));
        fx = dyte * (up * t1 + un * te) * m0 * mk(k, km_e);
      }
    }
    const double fxw = lane_prev(fx);
    // y flux through the north face (:2573-2587, 2623-2635)
    if (yface) {
      const double v = J.V[oy + P2 * k0];
      const double vel = dxtn * v;
      const double upos = 0.5 * (vel + fabs(vel));
      const double uneg = 0.5 * (vel - fabs(vel));
      const double rnormsk = m0 * (1.0 - mk(k, km_s));
      const double soutmsk = mk(k, km_n) * (1.0 - mk(k, km_n2));
      double fy = vel * (qy0 * t1 + qy1 * tn) -
                  upos * (cyp0 * tn + cyp1 * t1 + cyp2 * (ts * (1.0 - rnormsk) + t1 * rnormsk)) -
                  uneg * (cyn0 * (tn2 * (1.0 - soutmsk) + tn * soutmsk) + cyn1 * tn + cyn2 * t1);
      if (J.lwu && L == 1.0) {
        const double up = 0.5 * (v + fabs(v)), un = 0.5 * (v - fabs(v));
        fy = dxtn * (up * t1 + un * tn) * m0 * mk(k, km_n);
      }
      s_fy[row][lane] = fy;
    }
    __syncthreads();

    if (cell) {
      const int o = off + P2 * k0;
      const double fy = s_fy[row][lane], fys = s_fy[row - 1][lane];
      const double res_h = m0 * (fx - fxw + fy - fys) * datr;
      // vertical flux through the bottom face (:3002-3028)
      const double vel = J.W[off + P2 * k];
      const double upos = 0.5 * (vel + fabs(vel));
      const double uneg = 0.5 * (vel - fabs(vel));
      const double mkp1 = mk(kp1, km);
      double ft2;
      if (L == 1.0) {
        ft2 = (uneg * t1 + upos * t2) * m0 * mkp1;
      } else {
        // kp2 = p2 where that level is wet, else kp1 (:3010)
        const double tkp2 = (mk(p2, km) == 1.0) ? t3 : t2;
        const double upmsk = m0 * (1.0 - mk(km1, km));
        ft2 = vel * (quick_z[k - 1] * t1 + quick_z[k - 1 + nk] * t2) -
              uneg * (czp[k - 1] * t2 + czp[k - 1 + nk] * t1 + czp[k - 1 + 2 * nk] * t0) -
              upos * (czn[k - 1] * tkp2 + czn[k - 1 + nk] * t2 +
                      czn[k - 1 + 2 * nk] * (t1 * (1.0 - upmsk) + t2 * upmsk));
      }
      const double res_v = m0 * (ft1 - ft2);
      ft1 = ft2;
      double th = J.th ? J.th[o] : 0.0;
      th = th + (-res_h);
      th = th + (-res_v);
      const double rr = J.Rr ? J.Rr[o] : 1.0 / (J.R[o] + 1.0e-40);
      J.out[o] = (J.Rm1[o] * t1 + dtime * th) * rr;
    }
    // no barrier: every read of s_t precedes the second barrier above, and
    // the next level writes s_fy only after its first barrier

    // advance the k window
    t0 = t1; t1 = t2; t2 = t3;
    if (ok && k < nk) t3 = T[off + P2 * (min(k + 3, nk) - 1)];
    if constexpr (PART) {
      tp1 = tp2;
      if (ok && k < nk) tp2 = Tp[off + P2 * (min(k + 2, nk) - 1)];
    }
  }
}

template <int NW, bool PART>
int launch_one(mom_gpu_ctx *ctx, QStepJob J) {
  constexpr int TY = NW - 4;
  const KDom &d = ctx->kd;
  J.ntx = (d.ni + TX - 1) / TX;
  J.ntiles = J.ntx * ((d.nj + TY - 1) / TY);
  const int per = (J.ntiles + 7) / 8;
  hipLaunchKernelGGL((k_quick_step<NW, PART>), dim3(8 * per), dim3(RX * NW), 0, ctx->stream, J,
                     per);
  return MOM_GPU_OK;
}

}  // namespace

// The Quicker tracers of a batched step (scheme[n] == MOM_ADVECT_QUICKER,
// FP64 only): one pass each.  it/is: 0-based temp/salt indices (-1: none);
// with ts_same their limit flags are OR'ed (the partner's T(tau) is read).
int launch_quicker_step(mom_gpu_ctx *ctx, int ntr, const void *const *T_in, void *const *T_out,
                        const int *scheme, const double *const *th_in, const double *mn,
                        const double *mx, int it, int is, int ts_same, int lwu,
                        const double *u, const double *v, const double *w, const double *rho,
                        const double *rho_taum1, const double *rho_dztr, double dtime) {
  for (int n = 0; n < ntr; ++n) {
    if (scheme[n] != MOM_ADVECT_QUICKER) continue;
    QStepJob J{};
    J.d = ctx->kd;
    J.T = (const double *)T_in[n];
    J.out = (double *)T_out[n];
    J.th = th_in ? th_in[n] : nullptr;
    J.U = u; J.V = v; J.W = w; J.R = rho; J.Rm1 = rho_taum1; J.Rr = rho_dztr;
    J.kmt = ctx->kmt;
    J.dyte = ctx->dyte; J.dxtn = ctx->dxtn; J.datr = ctx->datr;
    J.qw = ctx->qw; J.qz = ctx->qzw;
    J.dtime = dtime;
    J.lo = mn[n]; J.hi = mx[n];
    J.lwu = lwu ? 1 : 0;
    int partner = -1;
    if (ts_same && it >= 0 && is >= 0) {
      if (n == it) partner = is;
      else if (n == is) partner = it;
    }
    if (partner >= 0) {
      J.Tp = (const double *)T_in[partner];
      J.plo = mn[partner]; J.phi = mx[partner];
      launch_one<16, true>(ctx, J);
    } else {
      launch_one<16, false>(ctx, J);
    }
    MOM_LAUNCHED(ctx, "k_quick_step");
  }
  return MOM_GPU_OK;
}
