// Fused MDFL-Sweby / DST-linear tracer advection for gfx950.
//
// One work-group owns a tile of TX x TY columns and marches down the k
// levels.  At each level it
//   Z  computes the provisional tracer of the z-sweep on the tile plus a
//      two-cell halo (ocean_tracer_advect.F90:3843-3909); the T window
//      T(k-1..k+2), the flux through the cell top and w(k-1) stay in
//      registers from one level to the next;
//   X  forms the flux through each cell's eastern face from DPP lane
//      shifts (one region row = one 64-lane wavefront) and the x-sweep
//      update (:3916-3969) -- no LDS and no barrier;
//   Y  stages the x-swept tracer in LDS, forms the northern-face fluxes and
//      finishes with the y-sweep, the divergence-correction removal and the
//      tendency (:3974-4056), accumulated into th_tendency.
// The halo columns/rows are recomputed by each tile with the same
// arithmetic as the tile that owns them, which is what the two
// mpp_update_domains calls of the reference (:3912, :3970) deliver.
//
// FLAVOR 0 orders its updates like advect_tracer_mdfl_sweby (per-tracer
// routine, with a sweby_limiter argument); FLAVOR 1 like
// advect_tracer_sweby_all (:4104-4511).  Limited fluxes (sweby_limiter = 1)
// and the update factors come from mdfl64.h, the arithmetic the batched
// step shares; other sweby_limiter values (DST-linear: 0) keep the
// reference's quotient form of psi, whose 1/Rj singularity the unlimited
// scheme needs.
#include "mom_internal.h"
#include "mdfl64.h"

namespace {

constexpr int RX = 64;            // region width (one wavefront)
constexpr int TX = RX - 4;        // interior tile width
// wavefronts per work-group: 8 with 2 or 4 region rows a wavefront, 16 with 1
constexpr int nw_of(int rpw) { return rpw == 1 ? 16 : 8; }
constexpr double ONESIXTH = 1.0 / 6.0;   // ocean_parameters.F90:92
constexpr double EPS30 = 1.0e-30;

__device__ inline double mk(int k, int kmt) { return (k <= kmt) ? 1.0 : 0.0; }

template <int FLAVOR>
__device__ inline double psi(double d0, double d1, double theta, double cfl,
                             double lim) {
  if (FLAVOR == 0) {
    const double p = d0 + d1 * theta;
    return p * (1.0 - lim) +
           fmx(0.0, fmn(fmn(1.0, d0 + d1 * theta), (1.0 - cfl) / (EPS30 + cfl) * theta)) * lim;
  } else {
    return fmx(0.0, fmn(fmn(1.0, d0 + d1 * theta), (1.0 - cfl) / (EPS30 + cfl) * theta));
  }
}

// Flux through one face: Tpos is the tracer upstream for a positive mass
// flux, Tneg for a negative one; ma, mb the two masks in reference order.
template <int FLAVOR>
__device__ inline double face_flux(double Rjm, double Rj, double Rjp,
                                   double massflux, double cfl, double Tpos,
                                   double Tneg, double ma, double mb,
                                   double lim) {
  const double d0 = (2.0 - cfl) * (1.0 - cfl) * ONESIXTH;
  const double d1 = (1.0 - cfl * cfl) * ONESIXTH;
  const double thetaP = Rjm / (EPS30 + Rj);
  const double thetaM = Rjp / (EPS30 + Rj);
  const double psiP = psi<FLAVOR>(d0, d1, thetaP, cfl, lim);
  const double psiM = psi<FLAVOR>(d0, d1, thetaM, cfl, lim);
  return 0.5 * ((massflux + fabs(massflux)) * (Tpos + psiP * Rj) +
                (massflux - fabs(massflux)) * (Tneg - psiM * Rj)) * ma * mb;
}

// LIM1: sweby_limiter = 1 (the limited flux of mdfl64.h); otherwise the
// quotient form with the limiter weight lim
template <int FLAVOR, int RPW, bool LIM1>
__global__ __launch_bounds__(RX * nw_of(RPW)) void k_mdfl(
    KDom d, const double *__restrict__ T, const double *__restrict__ U,
    const double *__restrict__ V, const double *__restrict__ W,
    const double *__restrict__ R, const int *__restrict__ kmt,
    const double *__restrict__ dat, const double *__restrict__ datr,
    const double *__restrict__ dxte, const double *__restrict__ dyte,
    const double *__restrict__ dxtn, const double *__restrict__ dytn,
    double dtime, double lim, double *__restrict__ th,
    double *__restrict__ wrk1, double *__restrict__ fxo,
    double *__restrict__ fyo, double *__restrict__ fzo, int ntx, int ntiles,
    int per_xcd) {
  constexpr int NW = nw_of(RPW);
  constexpr int RY = NW * RPW;
  constexpr int TY = RY - 4;
  __shared__ double s_tx[RY][RX];
  __shared__ double s_fy[RY][RX];
  __shared__ int s_km[RY][RX];

  // Blocks are dealt to the 8 XCDs round-robin; give each XCD a contiguous
  // run of positions of the order of 8-wide strips (strip_tile), so that the
  // halo rows and columns that neighbouring tiles both read hit its own L2.
  const int b = blockIdx.x;
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= ntiles) return;
  const int tile = strip_tile<8>(t, ntx, ntiles / ntx);
  const int tix = tile % ntx, tiy = tile / ntx;
  const int gi0 = tix * TX, gj0 = tiy * TY;

  const int lane = threadIdx.x & (RX - 1);
  const int wave = threadIdx.x / RX;
  const int nk = d.nk, P2 = d.P2;
  const int gi = gi0 - 2 + lane;
  constexpr bool limited = LIM1;
  const double rdt = 1.0 / dtime;

  int si = gi;
  bool colok = true;
  if (si < 0 || si >= d.ni) {
    if (d.cyclic_x) si = wrapi(si, d.ni); else colok = false;
  }

  int km[RPW], offz[RPW], gjr[RPW], sjr[RPW];
  bool ok[RPW], rowok[RPW];
  double t0[RPW], t1[RPW], t2[RPW], t3[RPW], t4[RPW], ftp[RPW], wkm1[RPW];
  double datv[RPW], datrv[RPW];
#pragma unroll
  for (int q = 0; q < RPW; ++q) {
    const int ry = wave + NW * q;
    const int gj = gj0 - 2 + ry;
    int sj = gj;
    bool rok = true;
    if (sj < 0 || sj >= d.nj) {
      if (d.cyclic_y) sj = wrapi(sj, d.nj); else rok = false;
    }
    gjr[q] = gj;
    sjr[q] = sj;
    rowok[q] = rok;
    ok[q] = colok && rok;
    offz[q] = ok[q] ? d2(d, si, sj) : 0;
    km[q] = ok[q] ? kmt[offz[q]] : 0;
    s_km[ry][lane] = km[q];
    datv[q] = ok[q] ? dat[offz[q]] : 0.0;
    datrv[q] = ok[q] ? datr[offz[q]] : 0.0;
    ftp[q] = 0.0;
    wkm1[q] = 0.0;
    if (ok[q]) {
      const int o = offz[q];
      t0[q] = T[o];
      t1[q] = t0[q];
      t2[q] = T[o + P2 * (min(2, nk) - 1)];
      t3[q] = T[o + P2 * (min(3, nk) - 1)];
    } else {
      t0[q] = t1[q] = t2[q] = t3[q] = 0.0;
    }
    t4[q] = 0.0;
  }
  __syncthreads();

  // The faces' offsets (0 where a face does not exist) and metrics, and a
  // level's 3-D loads issued when the level starts (u and the east face's two
  // rho_dzt, v and the north face's two, th_tendency of the point), so that
  // the x, y and final phases, each after a barrier, do not each wait for
  // their own loads with every wavefront of the work-group at once.  The
  // loads are unconditional (a value of a face that does not exist is
  // unused), so that no branch merge waits for them.
  int oxq[RPW], oyq[RPW];
  double mdyte[RPW], mdxte[RPW], mdxtn[RPW], mdytn[RPW];
#pragma unroll
  for (int q = 0; q < RPW; ++q) {
    const int ry = wave + NW * q;
    const int gj = gjr[q];
    const bool fx = rowok[q] && lane >= 1 && lane <= RX - 3 && gi >= -1 && gi <= d.ni - 1;
    const bool fy = ry >= 1 && ry <= RY - 3 && lane >= 2 && lane < RX - 2 && gi < d.ni &&
                    gj >= -1 && gj <= d.nj - 1;
    oxq[q] = fx ? d2(d, gi, sjr[q]) : 0;
    oyq[q] = fy ? d2(d, gi, gj) : 0;
    mdyte[q] = dyte[oxq[q]];
    mdxte[q] = dxte[oxq[q]];
    mdxtn[q] = dxtn[oyq[q]];
    mdytn[q] = dytn[oyq[q]];
  }

  // The 1-row shape (registers to spare) issues a level's w, rho_dzt and east
  // face loads (u and its two rho_dzt) during the level before, after its z
  // sweep: the sweeps that start a level then wait for nothing.
  constexpr bool PRE = RPW == 1;
  double nw[RPW], nr[RPW], nu[RPW], nrw[RPW], nre[RPW];
  if constexpr (PRE) {
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      nw[q] = W[offz[q] + P2];
      nr[q] = R[offz[q]];
      nu[q] = U[oxq[q]];
      nrw[q] = R[oxq[q]];
      nre[q] = R[oxq[q] + 1];
    }
  }
  for (int k = 1; k <= nk; ++k) {
    const int k0 = k - 1;
    const int kp1 = min(k + 1, nk), kp2 = min(k + 2, nk), km1 = max(k - 1, 1);
    double tz[RPW], rhov[RPW], dtrv[RPW], wv[RPW], fbtv[RPW], txv[RPW], mfe[RPW], mfw[RPW];
    double fxv[RPW], xfv[RPW];
    // (4-row shapes have no registers for them and load in the phases)
    constexpr bool EARLY = RPW <= 2;
    const size_t pk = (size_t)P2 * k0;
    double lu[RPW], lrw[RPW], lre[RPW], lv[RPW], lrs[RPW], lrn[RPW], lth[RPW], lw[RPW], lr[RPW];
    if constexpr (EARLY) {
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        const int ox = oxq[q], oy = oyq[q];
        if constexpr (PRE) {
          lw[q] = nw[q]; lr[q] = nr[q];
          lu[q] = nu[q]; lrw[q] = nrw[q]; lre[q] = nre[q];
        } else {
          lw[q] = W[offz[q] + P2 * k];
          lr[q] = R[offz[q] + pk];
          lu[q] = U[ox + pk];
          lrw[q] = R[ox + pk];
          lre[q] = R[ox + 1 + pk];
        }
        lv[q] = V[oy + pk];
        lrs[q] = R[oy + pk];
        lrn[q] = R[oy + d.NI + pk];
        lth[q] = th[offz[q] + pk];
      }
    }

    // ---- Z: flux through the bottom face and the z-sweep update ----------
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      tz[q] = 0.0; rhov[q] = 1.0; dtrv[q] = 0.0; wv[q] = 0.0; fbtv[q] = 0.0;
      if (ok[q]) {
        const int o = offz[q];
        const double w = EARLY ? lw[q] : W[o + P2 * k];
        const double r = EARLY ? lr[q] : R[o + P2 * k0];
        const double mkm1 = mk(km1, km[q]), mk0 = mk(k, km[q]);
        const double mkp1 = mk(kp1, km[q]), mkp2 = mk(kp2, km[q]);
        const double Rjp = (t0[q] - t1[q]) * mkm1 * mk0;
        const double Rj = (t1[q] - t2[q]) * mk0 * mkp1;
        const double Rjm = (t2[q] - t3[q]) * mkp1 * mkp2;
        const double massflux = datv[q] * w;
        const double dtr = dtime * mdfl64::rcp64(r);
        const double cfl = mdfl64::cfl_z(w, dtr);
        const double fbt =
            limited ? mdfl64::flux(mdfl64::face(massflux, cfl, mkp1 * mk0), Rjm, Rj, Rjp, t2[q],
                                   t1[q])
                    : face_flux<FLAVOR>(Rjm, Rj, Rjp, massflux, cfl, t2[q], t1[q], mkp1, mk0,
                                        lim);
        if (FLAVOR == 0) {
          tz[q] = t1[q] + dtr * (datrv[q] * (fbt - ftp[q]) + t1[q] * (wkm1[q] - w));
        } else {
          const double wk = datrv[q] * (fbt - ftp[q]) + t1[q] * (wkm1[q] - w);
          tz[q] = t1[q] + wk * dtr;
        }
        rhov[q] = r; dtrv[q] = dtr; wv[q] = w; fbtv[q] = fbt;
      }
    }

    // the tracer of level k+3, issued now that the z sweep has used t3 (a
    // point outside the domain keeps 0; the last levels reload level nk)
    if constexpr (EARLY) {
#pragma unroll
      for (int q = 0; q < RPW; ++q)
        if (ok[q]) t4[q] = T[offz[q] + P2 * (min(k + 3, nk) - 1)];
    }
    if constexpr (PRE) {
      // level k+1's (the last level reloads its own)
      const size_t pw = (size_t)P2 * min(k + 1, nk), pr = (size_t)P2 * min(k, nk - 1);
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        nw[q] = W[offz[q] + pw];
        nr[q] = R[offz[q] + pr];
        nu[q] = U[oxq[q] + pr];
        nrw[q] = R[oxq[q] + pr];
        nre[q] = R[oxq[q] + 1 + pr];
      }
    }

    // ---- X: eastern-face flux (shuffles) and the x-sweep update ----------
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int ry = wave + NW * q;
      const double tzm1 = lane_prev(tz[q]);
      const double tzp1 = lane_next(tz[q]);
      const double tzp2 = lane_next(tzp1);
      const int kmm1 = dpp_prev(km[q]);
      const int kmp1 = dpp_next(km[q]);
      const int kmp2 = dpp_next(kmp1);
      double fx = 0.0, mf = 0.0;
      if (rowok[q] && lane >= 1 && lane <= RX - 3 && gi >= -1 && gi <= d.ni - 1) {
        const int ox = oxq[q];
        const double u = EARLY ? lu[q] : U[ox + pk];
        mf = mdyte[q] * u;
        const double rw = EARLY ? lrw[q] : R[ox + pk], re = EARLY ? lre[q] : R[ox + 1 + pk];
        const double cfl = mdfl64::cfl_h(u, dtime, rw + re, mdxte[q]);
        const double mi = mk(k, km[q]), mim1 = mk(k, kmm1);
        const double mip1 = mk(k, kmp1), mip2 = mk(k, kmp2);
        const double Rjp = (tzp2 - tzp1) * mip2 * mip1;
        const double Rj = (tzp1 - tz[q]) * mip1 * mi;
        const double Rjm = (tz[q] - tzm1) * mi * mim1;
        fx = limited ? mdfl64::flux(mdfl64::face(mf, cfl, mi * mip1), Rjm, Rj, Rjp, tz[q], tzp1)
                     : face_flux<FLAVOR>(Rjm, Rj, Rjp, mf, cfl, tz[q], tzp1, mi, mip1, lim);
      }
      const double fxw = lane_prev(fx);
      const double mw = lane_prev(mf);
      double tx = 0.0, xf = 0.0;
      if (rowok[q] && lane >= 2 && lane < RX - 2 && gi < d.ni) {
        const double mi = mk(k, km[q]);
        if (FLAVOR == 0) {
          xf = mdfl64::sweep_factor(dtrv[q], mi, datrv[q]);
          tx = tz[q] + xf * (fxw - fx + t1[q] * (mf - mw));
        } else {
          const double wk = mi * datrv[q] * (fxw - fx + t1[q] * (mf - mw));
          tx = tz[q] + wk * dtrv[q];
        }
      }
      s_tx[ry][lane] = tx;
      txv[q] = tx; xfv[q] = xf; mfe[q] = mf; mfw[q] = mw; fxv[q] = fx;
    }
    __syncthreads();

    // ---- Y: northern-face flux from the staged x-swept tracer -------------
    double fyv[RPW];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int ry = wave + NW * q;
      const int gj = gjr[q];
      double fy = 0.0;
      if (ry >= 1 && ry <= RY - 3 && lane >= 2 && lane < RX - 2 && gi < d.ni &&
          gj >= -1 && gj <= d.nj - 1) {
        const int oy = oyq[q];
        const double v = EARLY ? lv[q] : V[oy + pk];
        const double mf = mdxtn[q] * v;
        const double rs = EARLY ? lrs[q] : R[oy + pk];
        const double rn = EARLY ? lrn[q] : R[oy + d.NI + pk];
        const double cfl = mdfl64::cfl_h(v, dtime, rs + rn, mdytn[q]);
        const double tjm1 = s_tx[ry - 1][lane], tj = s_tx[ry][lane];
        const double tjp1 = s_tx[ry + 1][lane], tjp2 = s_tx[ry + 2][lane];
        const double mjm1 = mk(k, s_km[ry - 1][lane]), mj = mk(k, s_km[ry][lane]);
        const double mjp1 = mk(k, s_km[ry + 1][lane]), mjp2 = mk(k, s_km[ry + 2][lane]);
        const double Rjp = (tjp2 - tjp1) * mjp2 * mjp1;
        const double Rj = (tjp1 - tj) * mjp1 * mj;
        const double Rjm = (tj - tjm1) * mj * mjm1;
        fy = limited ? mdfl64::flux(mdfl64::face(mf, cfl, mj * mjp1), Rjm, Rj, Rjp, tj, tjp1)
                     : face_flux<FLAVOR>(Rjm, Rj, Rjp, mf, cfl, tj, tjp1, mj, mjp1, lim);
      }
      s_fy[ry][lane] = fy;
      fyv[q] = fy;
    }
    __syncthreads();

    // ---- final: y-sweep update, correction removal, tendency -------------
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int ry = wave + NW * q;
      const int gj = gjr[q];
      const bool cell = ry >= 2 && ry < RY - 2 && lane >= 2 && lane < RX - 2 &&
                        gi < d.ni && gj < d.nj;
      if (cell) {
        const double fys = s_fy[ry - 1][lane], fyn = fyv[q];
        const double mi = mk(k, km[q]);
        const double r = rhov[q], Tc = t1[q];
        double w1;
        if (FLAVOR == 0) {
          double tr = txv[q] + xfv[q] * (fys - fyn);
          tr = tr + dtrv[q] * Tc * (wv[q] - wkm1[q] + datrv[q] * (mfw[q] - mfe[q]));
          w1 = mdfl64::tendency(r, tr, Tc, rdt, mi);
        } else {
          const double wk = mi * datrv[q] * (fys - fyn) +
                            Tc * (wv[q] - wkm1[q] + datrv[q] * (mfw[q] - mfe[q]));
          const double tr = txv[q] + wk * dtrv[q];
          w1 = mdfl64::tendency(r, tr, Tc, rdt, mi);
        }
        const int o = offz[q] + P2 * k0;
        th[o] = (EARLY ? lth[q] : th[o]) + w1;
        if (wrk1) wrk1[o] = w1;
        if (fzo) fzo[o] = fbtv[q];
        if (fxo) fxo[o] = fxv[q];
        if (fyo) fyo[o] = fyn;
      }
      // the western/southern boundary faces belong to the first tile
      if (fxo && tix == 0 && lane == 1 && ry >= 2 && ry < RY - 2 && gj < d.nj)
        fxo[d2(d, -1, gj) + P2 * k0] = fxv[q];
      if (fyo && tiy == 0 && ry == 1 && lane >= 2 && lane < RX - 2 && gi < d.ni)
        fyo[d2(d, gi, -1) + P2 * k0] = fyv[q];
    }

    // ---- advance the k window -------------------------------------------
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      ftp[q] = fbtv[q];
      wkm1[q] = wv[q];
      t0[q] = t1[q]; t1[q] = t2[q]; t2[q] = t3[q];
      if constexpr (EARLY) t3[q] = t4[q];
      else if (ok[q] && k < nk) t3[q] = T[offz[q] + P2 * (min(k + 3, nk) - 1)];
    }
  }
}

}  // namespace

template <int FLAVOR, int RPW, bool LIM1>
static void launch_mdfl_t(mom_gpu_ctx *ctx, const double *T, const double *u, const double *v,
                          const double *w, const double *rho, double dtime, double lim,
                          double *th, double *wrk1, double *fx, double *fy, double *fz) {
  constexpr int NW = nw_of(RPW);
  constexpr int TY = NW * RPW - 4;
  const KDom &d = ctx->kd;
  const int ntx = (d.ni + TX - 1) / TX;
  const int nty = (d.nj + TY - 1) / TY;
  const int ntiles = ntx * nty;
  const int per = (ntiles + 7) / 8;
  hipLaunchKernelGGL((k_mdfl<FLAVOR, RPW, LIM1>), dim3(8 * per), dim3(RX * NW), 0, ctx->stream,
                     d, T,
                     u, v, w, rho, ctx->kmt, ctx->dat, ctx->datr, ctx->dxte, ctx->dyte,
                     ctx->dxtn, ctx->dytn, dtime, lim, th, wrk1, fx, fy, fz, ntx, ntiles, per);
}

int launch_mdfl(mom_gpu_ctx *ctx, int flavor, const double *T,
                const double *u, const double *v, const double *w,
                const double *rho, double dtime, double lim, double *th,
                double *wrk1, double *fx, double *fy, double *fz) {
  const int rpw = ctx->tune.mdfl_rows_per_wave;
  const bool lim1 = lim == 1.0;
  if (flavor == 0 && lim1 && rpw == 1)
    launch_mdfl_t<0, 1, true>(ctx, T, u, v, w, rho, dtime, lim, th, wrk1, fx, fy, fz);
  else if (flavor == 0 && rpw == 1)
    launch_mdfl_t<0, 1, false>(ctx, T, u, v, w, rho, dtime, lim, th, wrk1, fx, fy, fz);
  else if (flavor == 1 && rpw == 1)
    launch_mdfl_t<1, 1, true>(ctx, T, u, v, w, rho, dtime, lim, th, wrk1, fx, fy, fz);
  else if (flavor == 0 && lim1 && rpw == 2)
    launch_mdfl_t<0, 2, true>(ctx, T, u, v, w, rho, dtime, lim, th, wrk1, fx, fy, fz);
  else if (flavor == 0 && lim1)
    launch_mdfl_t<0, 4, true>(ctx, T, u, v, w, rho, dtime, lim, th, wrk1, fx, fy, fz);
  else if (flavor == 0 && rpw == 2)
    launch_mdfl_t<0, 2, false>(ctx, T, u, v, w, rho, dtime, lim, th, wrk1, fx, fy, fz);
  else if (flavor == 0)
    launch_mdfl_t<0, 4, false>(ctx, T, u, v, w, rho, dtime, lim, th, wrk1, fx, fy, fz);
  else if (rpw == 2)
    launch_mdfl_t<1, 2, true>(ctx, T, u, v, w, rho, dtime, lim, th, wrk1, fx, fy, fz);
  else
    launch_mdfl_t<1, 4, true>(ctx, T, u, v, w, rho, dtime, lim, th, wrk1, fx, fy, fz);
  MOM_LAUNCHED(ctx, "k_mdfl");
  return MOM_GPU_OK;
}
