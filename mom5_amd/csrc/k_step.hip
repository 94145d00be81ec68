// Batched tracer step: MDFL-Sweby advection of a group of tracers and the
// explicit update (ocean_tracer.F90:2341-2350) in one k-marching pass.
//
// The work-group layout is that of k_mdfl (64-wide region rows, one row per
// wavefront, 2-cell halo, x sweep through shuffles, y sweep through LDS).
// The tracer-independent terms of every face (mass flux, CFL and the
// limiter coefficients d0, d1, (1-c)/(eps+c)) and of every update (dt/rho,
// dt*tmask*datr/rho, the divergence brackets) are formed once per point and
// level and reused for the G tracers of the group.
//
// FP64 groups take every face flux and update factor from mdfl64.h, the
// arithmetic of the per-tracer kernel, in the same order, so their T(taup1)
// equals, bit for bit, the per-tracer path (horz_advect_tracer into a zeroed
// th_tendency, then the explicit update).  FP32 groups store and compute the
// tracer in single precision and use the same division-free form of the
// limited flux:
//   psi*Rj = max(0, min(Rj, d0*Rj + d1*Rjm, k*Rjm))   for Rj >= 0,
//   psi*Rj = min(0, max(Rj, d0*Rj + d1*Rjm, k*Rjm))   for Rj < 0,
// which is the limiter of ocean_tracer_advect.F90:3877-3885 multiplied
// through by Rj (the 1e-30 guard is below FP32 resolution).
#include "mom_internal.h"
#include "mdfl64.h"

#include <algorithm>
#include <type_traits>
#include <cstring>

namespace {

constexpr int RX = 64;
constexpr int TX = RX - 4;
constexpr double ONESIXTH = 1.0 / 6.0;
constexpr double EPS30 = 1.0e-30;
constexpr double EPSLN = 1.0e-40;   // shared/constants/constants.F90:123

template <typename TT>
__device__ inline TT mk(int k, int kmt) { return (k <= kmt) ? TT(1) : TT(0); }

// tracer-independent limiter coefficients of one face
template <typename TT>
struct FaceC {
  TT M;      // mass flux through the face
  TT d0, d1, kk;
};

// FP32 groups: the same coefficients in single precision, with the
// hardware reciprocal (about 1 ulp) in place of the divisions
__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
// FP32 code uses explicit fmaf: the same fused operations in every kernel
// that inlines it, so FP32 results never depend on the launch shape
// (contraction left to the compiler could differ between instantiations).
__device__ inline FaceC<float> face_coef_f(float massflux, float cfl) {
  FaceC<float> c;
  c.M = massflux;
  c.d0 = (2.0f - cfl) * (1.0f - cfl) * (float)ONESIXTH;
  c.d1 = fmaf(-cfl, cfl, 1.0f) * (float)ONESIXTH;
  c.kk = (1.0f - cfl) * rcpf(1.0e-30f + cfl);
  return c;
}

// division-free limited flux (FP32).  The hardware median (v_med3_f32 on
// finite data) replaces compare-and-select: FP32 tracers need no particular
// signed zero, only the same code everywhere.
//   a >= 0: max(0, min(a, lin, kr)) = med3(0, a, min(lin, kr)),
//   a <  0: min(0, max(a, lin, kr)) = med3(0, a, max(lin, kr)),
// and min(lin, kr) = -med3(-lin, -kr, +inf), max(lin, kr) = -med3(-lin, -kr,
// -inf): sel = copysign(inf, a) picks the one the sign of a asks for (a = -0
// takes the max, whose median with 0 and a is a zero all the same), and the
// negations are source modifiers of the medians, so a limited term costs a
// bit insert and two medians.
__device__ inline float minmax_sel(float a) { return __builtin_copysignf(__builtin_inff(), a); }
// psi*a for the neighbouring difference r; d0a = d0*a
__device__ inline float psi_r(const FaceC<float> &c, float d0a, float a, float sel, float r) {
  const float lin = fmaf(c.d1, r, d0a);
  const float kr = c.kk * r;
  const float nm = __builtin_amdgcn_fmed3f(-lin, -kr, sel);
  return __builtin_amdgcn_fmed3f(0.0f, a, -nm);
}
// The upwind weight of a face with the face mask folded in.  Of the terms
// (M + |M|)/2*(Tpos + psiP) and (M - |M|)/2*(Tneg - psiM) only the upwind one
// has a nonzero weight, exactly M: the flux is M*mab*(Tu +- psi) with the
// upwind side's value and limiter, the same value as the two-term sum (up
// to the sign of a zero).
struct FaceW {
  float wm;    // M*mab
  float sg;    // +1 where M >= 0, -1 where M < 0
  bool pos;    // M >= 0
};
__device__ inline FaceW face_w(const FaceC<float> &c, float mab) {
  const bool pos = c.M >= 0.0f;
  return FaceW{c.M * mab, pos ? 1.0f : -1.0f, pos};
}
__device__ inline float flux_fast(const FaceC<float> &c, const FaceW &w, float Rjm, float Rj,
                                  float Rjp, float Tpos, float Tneg) {
  const float d0a = c.d0 * Rj;
  const float sel = minmax_sel(Rj);
  const float Rn = w.pos ? Rjm : Rjp;
  const float Tu = w.pos ? Tpos : Tneg;
  return w.wm * fmaf(w.sg, psi_r(c, d0a, Rj, sel, Rn), Tu);
}

// Level-sliced addressing: a buffer descriptor of the level (the array plus
// the level's offset, in scalar registers) and a 32-bit byte offset of the
// point within the level, so that no lane carries a 64-bit address.  A level
// of the largest domain is far below 4 GiB.
__device__ inline __amdgpu_buffer_rsrc_t lvl_rsrc(const void *p) {
  // raw buffer: stride 0, 4 GiB of records, 32-bit data format (CDNA)
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, -1, 0x00020000);
}
__device__ inline float ld_lv(const float *base, unsigned boff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(lvl_rsrc(base), (int)boff, 0, 0));
}
__device__ inline double ld_lv(const double *base, unsigned boff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(lvl_rsrc(base), (int)boff, 0, 0));
}
__device__ inline void st_lv(float *base, unsigned boff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), lvl_rsrc(base), (int)boff, 0, 0);
}
__device__ inline void st_lv(double *base, unsigned boff, double v) {
  using V2 = unsigned __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), lvl_rsrc(base), (int)boff, 0, 0);
}

template <typename TT>
struct StepTab {
  const TT *T[12];      // tracer at taum1 (= tau)
  TT *out[12];          // tracer at taup1
  const double *th[12]; // other th_tendency contributions (NULL: none)
};

// Everything one pass of one domain needs: the batched kernel takes one
// job by value; the ensemble kernel reads its jobs from a device table.
template <typename TT>
struct StepJob {
  KDom d;
  StepTab<TT> tab;
  const double *U, *V, *W, *R, *Rm1, *Rr;
  const int *kmt;
  const double *dat, *datr, *dxte, *dyte, *dxtn, *dytn;
  double dtime;
  int ntx, ntiles;
};

// One 64-wide tile of one job, marched through all levels.  Every thread of
// the block calls it with the same job and tile.
template <typename TT, int G, int RPW, int NW>
__device__ __forceinline__ void step_tile(const StepJob<TT> &J, int tile) {
  const KDom d = J.d;
  const StepTab<TT> &tab = J.tab;
  const double *__restrict__ U = J.U;
  const double *__restrict__ V = J.V;
  const double *__restrict__ W = J.W;
  const double *__restrict__ R = J.R;
  const double *__restrict__ Rm1 = J.Rm1;
  const double *__restrict__ Rr = J.Rr;
  const int *__restrict__ kmt = J.kmt;
  const double *__restrict__ dat = J.dat;
  const double *__restrict__ datr = J.datr;
  const double *__restrict__ dxte = J.dxte;
  const double *__restrict__ dyte = J.dyte;
  const double *__restrict__ dxtn = J.dxtn;
  const double *__restrict__ dytn = J.dytn;
  const double dtime = J.dtime;
  const double rdt = 1.0 / dtime;
  const int ntx = J.ntx;
  constexpr bool EXACT = sizeof(TT) == 8;
  constexpr int RY = NW * RPW;
  // FP64 groups of one row a wavefront and at most two tracers (registers to
  // spare) issue the tracers of level k+3 right after the z sweep (EARLYT;
  // the others load them at the end of the level), keep two copies of the
  // x-swept values and of the y fluxes, by level parity, where the LDS holds
  // them (ONEB: one barrier a level, the update of a level done after the
  // next level's barrier), and issue a level's w and rho_dzt during the level
  // before (PREWR).  FP32 groups of one row a wavefront take PREWR only, and
  // with it the loads of the east face (u and its two rho_dzt).
  constexpr bool EARLYT = RPW == 1 && EXACT && G <= 2;
  constexpr bool ONEB = EARLYT &&
                        2 * 2 * G * RY * RX * sizeof(TT) + RY * RX * sizeof(int) <= 160 * 1024;
  constexpr bool PREWR = EXACT ? ONEB : RPW == 1;
  constexpr int NB = ONEB ? 2 : 1;
  __shared__ TT s_tx[NB][G][RY][RX];
  __shared__ TT s_fy[NB][G][RY][RX];
  __shared__ int s_km[RY][RX];

  constexpr int TY = RY - 4;
  const int tix = tile % ntx, tiy = tile / ntx;
  const int gi0 = tix * TX, gj0 = tiy * TY;
  const int lane = threadIdx.x & (RX - 1);
  // the row of a wavefront, in a scalar register
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / RX);
  const int nk = d.nk, P2 = d.P2;
  const int gi = gi0 - 2 + lane;

  int si = gi;
  bool colok = true;
  if (si < 0 || si >= d.ni) {
    if (d.cyclic_x) si = wrapi(si, d.ni); else colok = false;
  }
  int km[RPW], offz[RPW], gjr[RPW], sjr[RPW];
  bool ok[RPW], rowok[RPW];
  // per-row state: FP64 for FP64 groups (reference rounding), FP32 for FP32
  // groups (fewer registers; their conservation weight reloads datr in FP64)
  using RT = typename std::conditional<EXACT, double, float>::type;
  RT datv[RPW], datrv[RPW], wkm1[RPW];
  TT t0[G][RPW], t1[G][RPW], t2[G][RPW], t3[G][RPW], t4[G][RPW], ftp[G][RPW];
#pragma unroll
  for (int q = 0; q < RPW; ++q) {
    const int ry = wave + NW * q;
    const int gj = gj0 - 2 + ry;
    int sj = gj;
    bool rok = true;
    if (sj < 0 || sj >= d.nj) {
      if (d.cyclic_y) sj = wrapi(sj, d.nj); else rok = false;
    }
    gjr[q] = gj; sjr[q] = sj; rowok[q] = rok;
    ok[q] = colok && rok;
    offz[q] = ok[q] ? d2(d, si, sj) : 0;
    km[q] = ok[q] ? kmt[offz[q]] : 0;
    s_km[ry][lane] = km[q];
    datv[q] = ok[q] ? (RT)dat[offz[q]] : RT(0);
    datrv[q] = ok[q] ? (RT)datr[offz[q]] : RT(0);
    wkm1[q] = RT(0);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      ftp[g][q] = TT(0);
      if (ok[q]) {
        const TT *T = tab.T[g];
        const int o = offz[q];
        t0[g][q] = T[o];
        t1[g][q] = t0[g][q];
        t2[g][q] = T[o + P2 * (min(2, nk) - 1)];
        t3[g][q] = T[o + P2 * (min(3, nk) - 1)];
      } else {
        t0[g][q] = t1[g][q] = t2[g][q] = t3[g][q] = TT(0);
      }
      t4[g][q] = TT(0);
    }
  }
  __syncthreads();

  // Mask depths of the x and y faces: for masks in {0, 1}, mk(k, a)*mk(k, b)
  // == mk(k, min(a, b)) bit for bit.  FP32 groups (register room to spare)
  // keep them across the k loop; FP64 groups re-form them per level.
  auto face_depths = [&](int q, int &x0, int &x1, int &x2, int &y0, int &y1, int &y2) {
    const int ry = wave + NW * q;
    const int kmm1 = lane_prev(km[q]);
    const int kmp1 = lane_next(km[q]);
    const int kmp2 = lane_next(kmp1);
    x0 = min(km[q], kmm1);
    x1 = min(kmp1, km[q]);
    x2 = min(kmp2, kmp1);
    const int kjm1 = s_km[max(ry - 1, 0)][lane];
    const int kjp1 = s_km[min(ry + 1, RY - 1)][lane];
    const int kjp2 = s_km[min(ry + 2, RY - 1)][lane];
    y0 = min(km[q], kjm1);
    y1 = min(kjp1, km[q]);
    y2 = min(kjp2, kjp1);
  };
  int kx0[RPW], kx1[RPW], kx2[RPW], ky0[RPW], ky1[RPW], ky2[RPW];
  // FP32 groups also keep the face metrics (dyte, dxte of the x face, dxtn,
  // dytn of the y face, in FP32 as they are used) and datr (FP64) in registers
  float mdyte[RPW], mdxte[RPW], mdxtn[RPW], mdytn[RPW];
  double mdatr[RPW];
  if constexpr (!EXACT) {
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      face_depths(q, kx0[q], kx1[q], kx2[q], ky0[q], ky1[q], ky2[q]);
      const int ry = wave + NW * q;
      const bool fx = rowok[q] && lane >= 1 && lane <= RX - 3 && gi >= -1 && gi <= d.ni - 1;
      const bool fy = ry >= 1 && ry <= RY - 3 && lane >= 2 && lane < RX - 2 && gi < d.ni &&
                      gjr[q] >= -1 && gjr[q] <= d.nj - 1;
      const int ox = d2(d, gi, sjr[q]), oy = d2(d, gi, gjr[q]);
      mdyte[q] = fx ? (float)dyte[ox] : 0.0f;
      mdxte[q] = fx ? (float)dxte[ox] : 0.0f;
      mdxtn[q] = fy ? (float)dxtn[oy] : 0.0f;
      mdytn[q] = fy ? (float)dytn[oy] : 0.0f;
      mdatr[q] = ok[q] ? datr[offz[q]] : 0.0;
    }
  }

  // FP64 groups issue all of a level's 3-D loads when the level starts (w
  // and rho_dzt of the point during the level before, PREWR): u and the two
  // rho_dzt of its east face, v and the two of its north face,
  // rho_dzt(taum1) and 1/rho_dzt(taup1) (or rho_dzt(taup1); with ONEB those
  // of the level before, for its deferred update), so that the x, y and
  // update phases (each after a barrier) do not each wait for their own loads
  // with every wavefront of the work-group at once.  The loads are
  // unconditional, from offset 0 where the face or point does not exist (the
  // value is then unused).  The face metrics stay in registers as for FP32
  // groups.
  struct LvIn {
    double w, r, u, rw, re, v, rs, rn, rm1, rr;
  };
  int oxq[RPW], oyq[RPW];
  double hdyte[RPW], hdxte[RPW], hdxtn[RPW], hdytn[RPW];
#pragma unroll
  for (int q = 0; q < RPW; ++q) {
    const int ry = wave + NW * q;
    const bool fx = rowok[q] && lane >= 1 && lane <= RX - 3 && gi >= -1 && gi <= d.ni - 1;
    const bool fy = ry >= 1 && ry <= RY - 3 && lane >= 2 && lane < RX - 2 && gi < d.ni &&
                    gjr[q] >= -1 && gjr[q] <= d.nj - 1;
    oxq[q] = fx ? d2(d, gi, sjr[q]) : 0;
    oyq[q] = fy ? d2(d, gi, gjr[q]) : 0;
  }
  if constexpr (EXACT) {
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      hdyte[q] = dyte[oxq[q]];
      hdxte[q] = dxte[oxq[q]];
      hdxtn[q] = dxtn[oyq[q]];
      hdytn[q] = dytn[oyq[q]];
    }
  }
  // byte offsets within a level: of the point (FP64 fields and tracers), of
  // its east face and of its north face (FP64 fields)
  unsigned bz[RPW], bt[RPW], bx[RPW], by[RPW];
#pragma unroll
  for (int q = 0; q < RPW; ++q) {
    bz[q] = (unsigned)offz[q] * 8u;
    bt[q] = (unsigned)offz[q] * (unsigned)sizeof(TT);
    bx[q] = (unsigned)oxq[q] * 8u;
    by[q] = (unsigned)oyq[q] * 8u;
  }
  const double *Rx = Rr ? Rr : R;
  // PREWR shapes issue a level's w and rho_dzt (FP32: and the east face's
  // u and rho_dzt) during the level before, after its z sweep: the sweeps
  // that start a level then wait for nothing
  double nw[RPW], nr[RPW], nu[RPW], nrw[RPW], nre[RPW];
  if constexpr (PREWR) {
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      nw[q] = ld_lv(W + P2, bz[q]);
      nr[q] = ld_lv(R, bz[q]);
      if constexpr (!EXACT) {
        nu[q] = ld_lv(U, bx[q]);
        nrw[q] = ld_lv(R, bx[q]);
        nre[q] = ld_lv(R + 1, bx[q]);
      }
    }
  }
  auto lv_load = [&](int k, LvIn *L) {
    const size_t pk = (size_t)P2 * (k - 1);
    // the update's rho_dzt(taum1) and 1/rho_dzt(taup1): of the level before
    // when the update is deferred (unused at the first level)
    const size_t pu = ONEB ? (size_t)P2 * max(k - 2, 0) : pk;
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      if constexpr (PREWR) {
        L[q].w = nw[q];
        L[q].r = nr[q];
      } else {
        L[q].w = ld_lv(W + pk + P2, bz[q]);
        L[q].r = ld_lv(R + pk, bz[q]);
      }
      L[q].u = ld_lv(U + pk, bx[q]);
      L[q].rw = ld_lv(R + pk, bx[q]);
      L[q].re = ld_lv(R + pk + 1, bx[q]);
      L[q].v = ld_lv(V + pk, by[q]);
      L[q].rs = ld_lv(R + pk, by[q]);
      L[q].rn = ld_lv(R + pk + d.NI, by[q]);
      L[q].rm1 = ld_lv(Rm1 + pu, bz[q]);
      L[q].rr = ld_lv(Rx + pu, bz[q]);
    }
  };

  // Deferred FP64 update of level kf (ONEB): the tracer at kf is t0 (the
  // window has moved on), the y fluxes are in copy kf & 1, the rest was kept
  // from level kf in pf_*.  The arithmetic is that of the in-level update.
  double pf_xf[RPW], pf_dtr[RPW], pf_r[RPW], pf_brk[RPW];
  TT pf_tx[G][RPW];
  auto update_deferred = [&](int kf, const LvIn *L) {
    const int bp = kf & 1;
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int ry = wave + NW * q;
      const int gj = gjr[q];
      const bool cell = ry >= 2 && ry < RY - 2 && lane >= 2 && lane < RX - 2 && gi < d.ni &&
                        gj < d.nj;
      if (cell) {
        const size_t pk = (size_t)P2 * (kf - 1);
        const double mi = mk<double>(kf, km[q]);
        const double rm1 = L[q].rm1;
        const double rr = Rr ? L[q].rr : 1.0 / (L[q].rr + EPSLN);
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const double thin = tab.th[g] ? ld_lv(tab.th[g] + pk, bz[q]) : 0.0;
          const double Tc = t0[g][q];
          double tr = pf_tx[g][q] + pf_xf[q] * (s_fy[bp][g][ry - 1][lane] - s_fy[bp][g][ry][lane]);
          tr = tr + pf_dtr[q] * Tc * pf_brk[q];
          const double th = thin + mdfl64::tendency(pf_r[q], tr, Tc, rdt, mi);
          st_lv(tab.out[g] + pk, bt[q], (rm1 * Tc + dtime * th) * rr);
        }
      }
    }
  };

  auto level = [&](const int k) {
    const int bb = ONEB ? (k & 1) : 0;
    LvIn cur[RPW];
    if constexpr (EXACT) {
      lv_load(k, cur);
    } else {
      // FP32 groups (registers for eight more values, not ten): the z, x and
      // y phases' loads
      const size_t pk = (size_t)P2 * (k - 1);
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        if constexpr (PREWR) {
          cur[q].w = nw[q];
          cur[q].r = nr[q];
        } else {
          cur[q].w = ld_lv(W + pk + P2, bz[q]);
          cur[q].r = ld_lv(R + pk, bz[q]);
        }
        if constexpr (PREWR) {
          cur[q].u = nu[q];
          cur[q].rw = nrw[q];
          cur[q].re = nre[q];
        } else {
          cur[q].u = ld_lv(U + pk, bx[q]);
          cur[q].rw = ld_lv(R + pk, bx[q]);
          cur[q].re = ld_lv(R + pk + 1, bx[q]);
        }
        cur[q].v = ld_lv(V + pk, by[q]);
        cur[q].rs = ld_lv(R + pk, by[q]);
        cur[q].rn = ld_lv(R + pk + d.NI, by[q]);
      }
    }
    const int k0 = k - 1;
    const int kp1 = min(k + 1, nk), kp2 = min(k + 2, nk);
    double rhov[RPW], xf64v[RPW], dtrv[RPW];   // dtrv: FP64 groups, dtime/rho_dzt
    RT wv[RPW];
    float rinv[RPW];   // FP32 groups: about 1/rho_dzt
    TT tz[G][RPW], fbtv[G][RPW], txv[G][RPW], fyv[G][RPW];
    TT dxf[G][RPW];    // FP32 groups: x flux difference of the cell
    TT mfe[RPW], mfw[RPW], xfac[RPW];

    // ---- Z ----------------------------------------------------------------
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      rhov[q] = 1.0; wv[q] = RT(0); rinv[q] = 1.0f; dtrv[q] = 0.0;
      // mk(km1)*mk(k) = mk(k), mk(k)*mk(kp1) = mk(kp1), mk(kp1)*mk(kp2) =
      // mk(kp2): km1 <= k <= kp1 <= kp2 and masks in {0, 1}
      const TT mA = mk<TT>(k, km[q]), mB = mk<TT>(kp1, km[q]), mC = mk<TT>(kp2, km[q]);
      if constexpr (!EXACT) {
        // FP32 groups take no branch: at a point outside the domain (t, dat,
        // datr and the masks 0) w = 0 and rho_dzt = 1 keep every value
        // finite, its tz 0 and its flux a zero
        const double w = ok[q] ? cur[q].w : 0.0;
        const double r = ok[q] ? cur[q].r : 1.0;
        rhov[q] = r; wv[q] = (RT)w;
        const float ri = rcpf((float)r);
        rinv[q] = ri;
        const float wf = (float)w;
        const FaceC<float> c = face_coef_f((float)datv[q] * wf, fabsf(wf * (float)dtime * ri));
        const FaceW wz = face_w(c, mB);
        const float dtrf = (float)dtime * ri, datrf = (float)datrv[q];
        const float wdiff = (float)wkm1[q] - wf;
#pragma unroll
        // where the face is wet (mB = 1) so is the cell above (mA = 1);
        // where it is dry its weights are 0: only mC needs applying
        for (int g = 0; g < G; ++g) {
          const float Rjp = t0[g][q] - t1[g][q];
          const float Rj = t1[g][q] - t2[g][q];
          const float Rjm = (t2[g][q] - t3[g][q]) * mC;
          const float fbt = flux_fast(c, wz, Rjm, Rj, Rjp, t2[g][q], t1[g][q]);
          tz[g][q] = fmaf(dtrf, fmaf(datrf, fbt - ftp[g][q], t1[g][q] * wdiff), t1[g][q]);
          fbtv[g][q] = fbt;
        }
      } else if (ok[q]) {
        const double w = cur[q].w, r = cur[q].r;
        rhov[q] = r; wv[q] = (RT)w;
        const double dtr = dtime * mdfl64::rcp64(r);
        dtrv[q] = dtr;
        const mdfl64::Face cz = mdfl64::face(datv[q] * w, mdfl64::cfl_z(w, dtr), mB);
        const double wdiff = wkm1[q] - w;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const double Rjp = (t0[g][q] - t1[g][q]) * mA;
          const double Rj = (t1[g][q] - t2[g][q]) * mB;
          const double Rjm = (t2[g][q] - t3[g][q]) * mC;
          const double fbt = mdfl64::flux(cz, Rjm, Rj, Rjp, t2[g][q], t1[g][q]);
          tz[g][q] = t1[g][q] + dtr * (datrv[q] * (fbt - ftp[g][q]) + t1[g][q] * wdiff);
          fbtv[g][q] = fbt;
        }
      } else {
#pragma unroll
        for (int g = 0; g < G; ++g) { tz[g][q] = TT(0); fbtv[g][q] = TT(0); }
      }
    }

    // The tracers of level k+3, issued now that the z sweep has used t3:
    // they have the x and y sweeps to arrive.  (A point outside the domain
    // keeps 0; the last levels reload level nk, unused.)
    if constexpr (EARLYT) {
      const size_t p3 = (size_t)P2 * (min(k + 3, nk) - 1);
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        if (ok[q]) {
#pragma unroll
          for (int g = 0; g < G; ++g) t4[g][q] = ld_lv(tab.T[g] + p3, bt[q]);
        }
      }
    }
    if constexpr (PREWR) {
      // w and rho_dzt of level k+1 (the last level reloads its own)
      const size_t pw = (size_t)P2 * min(k + 1, nk), pr = (size_t)P2 * min(k, nk - 1);
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        nw[q] = ld_lv(W + pw, bz[q]);
        nr[q] = ld_lv(R + pr, bz[q]);
        if constexpr (!EXACT) {
          nu[q] = ld_lv(U + pr, bx[q]);
          nrw[q] = ld_lv(R + pr, bx[q]);
          nre[q] = ld_lv(R + pr + 1, bx[q]);
        }
      }
    }

    // ---- X ----------------------------------------------------------------
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int ry = wave + NW * q;
      const bool face = rowok[q] && lane >= 1 && lane <= RX - 3 && gi >= -1 && gi <= d.ni - 1;
      if constexpr (EXACT) face_depths(q, kx0[q], kx1[q], kx2[q], ky0[q], ky1[q], ky2[q]);
      const TT mi = mk<TT>(k, km[q]);
      const TT mX1 = mk<TT>(k, kx1[q]);
      mdfl64::Face cx{};
      double Mx = 0.0;
      FaceC<float> cxf{0.0f, 0.0f, 0.0f, 0.0f};
      if constexpr (!EXACT) {
        // FP32 groups take no branch: where there is no face its metrics
        // are 0, so its mass flux is 0 and its CFL number is set to 0
        const float uf = (float)cur[q].u;
        const float cfl = fabsf(uf * (float)(dtime * 2.0) *
                                rcpf(((float)cur[q].rw + (float)cur[q].re) * mdxte[q]));
        cxf = face_coef_f(mdyte[q] * uf, face ? cfl : 0.0f);
      } else if (face) {
        const double u = cur[q].u, rw = cur[q].rw, re = cur[q].re;
        Mx = hdyte[q] * u;
        cx = mdfl64::face(Mx, mdfl64::cfl_h(u, dtime, rw + re, hdxte[q]), (double)mX1);
      }
      const bool cell = rowok[q] && lane >= 2 && lane < RX - 2 && gi < d.ni;
      double mw64 = 0.0;
      if constexpr (EXACT) {
        mw64 = lane_prev(Mx);
        // dtime*tmask*datr/rho (x and y sweeps)
        const double xf64 = mdfl64::sweep_factor(dtrv[q], (double)mi, datrv[q]);
        xf64v[q] = xf64;
        mfe[q] = (TT)Mx;
        mfw[q] = (TT)mw64;
        xfac[q] = (TT)xf64;
      } else {
        xf64v[q] = 0.0;
        mfe[q] = cxf.M;
        mfw[q] = lane_prev(cxf.M);
        xfac[q] = (float)dtime * mi * (float)datrv[q] * rinv[q];
      }
      FaceW wx{0.0f, 0.0f};
      if constexpr (!EXACT) wx = face_w(cxf, mX1);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        // Each lane forms its east face's masked difference once; the faces
        // either side have theirs one lane over: (T(i+2) - T(i+1))*mX2 is the
        // next lane's difference and (T(i) - T(i-1))*mX0 the previous lane's
        // (mX2, mX0 are the neighbouring lanes' mX1), the same values bit for bit.
        const TT tzp1 = lane_next(tz[g][q]);
        const TT Rjl = (tzp1 - tz[g][q]) * mX1;
        const TT Rjp = lane_next(Rjl);
        const TT Rjm = lane_prev(Rjl);
        // FP32: the unmasked difference would do (the face weights carry
        // mX1), the masked one gives the same flux; where there is no face
        // the weights are 0 and the flux a zero
        TT fx = TT(0);
        if constexpr (!EXACT)
          fx = flux_fast(cxf, wx, Rjm, Rjl, Rjp, tz[g][q], tzp1);
        else if (face)
          fx = mdfl64::flux(cx, Rjm, Rjl, Rjp, tz[g][q], tzp1);
        const TT fxw = lane_prev(fx);
        if constexpr (EXACT) {
          dxf[g][q] = fxw - fx;
        } else {
          // FP32: the z and x parts of the flux sum, formed now so that only
          // one value per tracer stays live through the y sweep
          dxf[g][q] = (fbtv[g][q] - ftp[g][q]) + (fxw - fx);
          ftp[g][q] = fbtv[g][q];
        }
        // FP32: formed everywhere (outside the domain xfac = 0 and tz = 0;
        // the halo columns' values are not read)
        TT tx = TT(0);
        if constexpr (!EXACT)
          tx = fmaf(xfac[q], fmaf(t1[g][q], mfe[q] - mfw[q], fxw - fx), tz[g][q]);
        else if (cell)
          tx = tz[g][q] + xf64v[q] * (fxw - fx + t1[g][q] * (Mx - mw64));
        s_tx[bb][g][ry][lane] = tx;
        txv[g][q] = tx;
      }
    }
    __syncthreads();
    if constexpr (ONEB) {
      if (k > 1) update_deferred(k - 1, cur);
    }

    // ---- Y ----------------------------------------------------------------
    // FP32 groups: the update's rho_dzt(taum1) and 1/rho_dzt(taup1) (or
    // rho_dzt(taup1)), issued now that the x phase's values are dead
    if constexpr (!EXACT) {
      const size_t pk = (size_t)P2 * k0;
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        cur[q].rm1 = ld_lv(Rm1 + pk, bz[q]);
        cur[q].rr = ld_lv(Rx + pk, bz[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int ry = wave + NW * q;
      const int gj = gjr[q];
      const TT mY2 = mk<TT>(k, ky2[q]), mY1 = mk<TT>(k, ky1[q]), mY0 = mk<TT>(k, ky0[q]);
      if constexpr (!EXACT) {
        // FP32 groups: a row (one wavefront) with no face skips the sweep;
        // within a row no branch, as in the x sweep.  The update reads the
        // faces of rows 1 .. RY-3 only.
        if (ry >= 1 && ry <= RY - 3 && gj >= -1 && gj <= d.nj - 1) {
          const bool face = lane >= 2 && lane < RX - 2 && gi < d.ni;
          const float vf = (float)cur[q].v;
          const float cfl = fabsf(vf * (float)(dtime * 2.0) *
                                  rcpf(((float)cur[q].rs + (float)cur[q].rn) * mdytn[q]));
          const FaceC<float> cyf = face_coef_f(mdxtn[q] * vf, face ? cfl : 0.0f);
          const FaceW wy = face_w(cyf, mY1);
#pragma unroll
          for (int g = 0; g < G; ++g) {
            const TT tjm1 = s_tx[bb][g][ry - 1][lane], tj = s_tx[bb][g][ry][lane];
            const TT tjp1 = s_tx[bb][g][ry + 1][lane], tjp2 = s_tx[bb][g][ry + 2][lane];
            const TT Rjp = (tjp2 - tjp1) * mY2;
            const TT Rj = tjp1 - tj;   // the face weights carry mY1
            const TT Rjm = (tj - tjm1) * mY0;
            s_fy[bb][g][ry][lane] = flux_fast(cyf, wy, Rjm, Rj, Rjp, tj, tjp1);
          }
        }
      } else {
        const bool face = ry >= 1 && ry <= RY - 3 && lane >= 2 && lane < RX - 2 && gi < d.ni &&
                          gj >= -1 && gj <= d.nj - 1;
        mdfl64::Face cy{};
        if (face) {
          const double v = cur[q].v, rs = cur[q].rs, rn = cur[q].rn;
          cy = mdfl64::face(hdxtn[q] * v, mdfl64::cfl_h(v, dtime, rs + rn, hdytn[q]),
                            (double)mY1);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
          TT fy = TT(0);
          if (face) {
            const TT tjm1 = s_tx[bb][g][ry - 1][lane], tj = s_tx[bb][g][ry][lane];
            const TT tjp1 = s_tx[bb][g][ry + 1][lane], tjp2 = s_tx[bb][g][ry + 2][lane];
            const TT Rjp = (tjp2 - tjp1) * mY2;
            const TT Rj = (tjp1 - tj) * mY1;
            const TT Rjm = (tj - tjm1) * mY0;
            fy = mdfl64::flux(cy, Rjm, Rj, Rjp, tj, tjp1);
          }
          s_fy[bb][g][ry][lane] = fy;
          fyv[g][q] = fy;
        }
      }
    }
    if constexpr (ONEB) {
      // keep what the deferred update of this level needs
#pragma unroll
      for (int q = 0; q < RPW; ++q) {
        pf_xf[q] = xf64v[q];
        pf_dtr[q] = dtrv[q];
        pf_r[q] = rhov[q];
        pf_brk[q] = wv[q] - wkm1[q] + datrv[q] * ((double)mfw[q] - (double)mfe[q]);
#pragma unroll
        for (int g = 0; g < G; ++g) pf_tx[g][q] = txv[g][q];
      }
    } else {
    __syncthreads();

    // ---- final: y update, correction removal, explicit update --------------
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int ry = wave + NW * q;
      const int gj = gjr[q];
      const bool cell = ry >= 2 && ry < RY - 2 && lane >= 2 && lane < RX - 2 && gi < d.ni &&
                        gj < d.nj;
      if (cell) {
        const size_t pk = (size_t)P2 * k0;
        const double r = rhov[q];
        const double mi = mk<double>(k, km[q]);
        const double brk = wv[q] - wkm1[q] + datrv[q] * ((double)mfw[q] - (double)mfe[q]);
        const double rm1 = cur[q].rm1;
        const double rr = Rr ? cur[q].rr : 1.0 / (cur[q].rr + EPSLN);
        // FP32: T(taup1) = T*(rho_taum1/rho_taup1) + dtime*datr/rho_taup1
        //                  * (sum of the face flux differences) + th*dtime/rho_taup1.
        // The update is rebuilt from the limited fluxes alone (flux form); the
        // T*div terms that the sweeps add for the limiter and the reference
        // removes at the end cancel in exact arithmetic but not in FP32.  The
        // factors are formed in FP64 without division, so every cell's content
        // weight rho*dat*factor is dtime to 1e-16 and the fluxes telescope:
        // what is left is the unbiased rounding of the FP32 result.
        const float a32 = (float)(rm1 * rr);
        double f64 = 0.0;
        if constexpr (!EXACT) f64 = dtime * rr * mdatr[q] * mi;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if constexpr (EXACT) {
            const double thin = tab.th[g] ? ld_lv(tab.th[g] + pk, bz[q]) : 0.0;
            const double Tc = t1[g][q];
            double tr = txv[g][q] + xf64v[q] * (s_fy[0][g][ry - 1][lane] - fyv[g][q]);
            tr = tr + dtrv[q] * Tc * brk;
            const double th = thin + mdfl64::tendency(r, tr, Tc, rdt, mi);
            st_lv(tab.out[g] + pk, bt[q], (rm1 * Tc + dtime * th) * rr);
          } else {
            const float Tc = t1[g][q];
            const float dflux = dxf[g][q] + (s_fy[0][g][ry - 1][lane] - s_fy[0][g][ry][lane]);
            // other th_tendency contributions: k_add_th_f32 adds
            // (float)th * c32 to this value afterwards, bit for bit the
            // same as here, so that FP32 passes hold no th pointers
            st_lv(tab.out[g] + pk, bt[q], Tc * a32 + (float)(f64 * (double)dflux));
          }
        }
      }
    }
    }   // !ONEB

    // ---- advance the k window -----------------------------------------------
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      wkm1[q] = wv[q];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if constexpr (EXACT) ftp[g][q] = fbtv[g][q];
        t0[g][q] = t1[g][q]; t1[g][q] = t2[g][q]; t2[g][q] = t3[g][q];
        if constexpr (EARLYT) t3[g][q] = t4[g][q];
      }
      if constexpr (!EARLYT) {
        const size_t p3 = (size_t)P2 * (min(k + 3, nk) - 1);
        if (ok[q]) {
#pragma unroll
          for (int g = 0; g < G; ++g) t3[g][q] = ld_lv(tab.T[g] + p3, bt[q]);
        }
      }
    }
  };
  for (int k = 1; k <= nk; ++k) level(k);
  if constexpr (ONEB) {
    LvIn last[RPW];
    const size_t pk = (size_t)P2 * (nk - 1);
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      last[q].rm1 = ld_lv(Rm1 + pk, bz[q]);
      last[q].rr = ld_lv(Rx + pk, bz[q]);
    }
    __syncthreads();
    update_deferred(nk, last);
  }
}


template <typename TT, int G, int RPW, int NW>
__global__ __launch_bounds__(RX * NW) void k_step(StepJob<TT> J, int per_xcd) {
  const int b = blockIdx.x;
  // XCD-aware: each XCD takes per_xcd consecutive positions of the order of
  // 8-wide strips, so x- and y-neighbours share their halo in one L2
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= J.ntiles) return;
  step_tile<TT, G, RPW, NW>(J, strip_tile<8>(t, J.ntx, J.ntiles / J.ntx));
}

// Ensemble pass: a persistent grid takes (job, tile) items from a device-side
// counter in order (jobs sorted largest first) until none is left.  first[j]
// is the index of job j's first item; total = sum of the jobs' tiles.  Every
// block leaves the loop once the counter passes total.
template <typename TT, int G, int RPW, int NW>
__global__ __launch_bounds__(RX * NW) void k_step_queue(const StepJob<TT> *__restrict__ jobs,
                                                        const int *__restrict__ first, int njobs,
                                                        int total, int *__restrict__ counter) {
  __shared__ int s_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= total) return;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (first[mid] <= item) lo = mid; else hi = mid - 1;
    }
    step_tile<TT, G, RPW, NW>(jobs[lo], item - first[lo]);
  }
}

// ---- launch shapes ------------------------------------------------------------
// A shape is (rows per wavefront, wavefronts per tile); each has the group
// sizes it is instantiated for.  Knob values that name no shape fall back to
// 8 wavefronts.  Tiles are RX x (RPW*NW) with a 2-cell halo.
//   FP64: (2, 8), (4, 8): 1, 2, 4 tracers; (2, 12): 1, 2 (168 VGPRs, 3 waves/SIMD);
//         (1, 16): 1, 2, 4 (at most 128 VGPRs, 4 waves/SIMD)
//   FP32: (2, 8), (4, 8), (2, 16): 1, 2, 4; (1, 16): 1, 2, 4, 6, 8, 10 (one row a
//         wavefront leaves room for ten tracers in 128 VGPRs); (2, 12): 1, 2, 5
struct Shape {
  int rpw, nw;
};

template <typename TT>
static Shape shape_of(int rpw, int waves) {
  if (sizeof(TT) == 4) {
    if (rpw == 1 && waves == 16) return {1, 16};
    if (rpw == 2 && waves == 16) return {2, 16};
    if (rpw == 2 && waves == 12) return {2, 12};
  } else {
    if (rpw == 2 && waves == 12) return {2, 12};
    if (rpw == 1 && waves == 16) return {1, 16};
  }
  return {rpw == 4 ? 4 : 2, 8};
}

// Group sizes of a shape, largest first (0-terminated).
template <typename TT>
static const int *group_sizes(Shape sh) {
  static const int g124[] = {4, 2, 1, 0}, g12[] = {2, 1, 0}, g125[] = {5, 2, 1, 0};
  static const int g12_4_8_10[] = {10, 8, 6, 4, 2, 1, 0};
  if (sizeof(TT) == 4) {
    if (sh.rpw == 1) return g12_4_8_10;
    if (sh.nw == 12) return g125;
    return g124;
  }
  return sh.nw == 12 ? g12 : g124;   // (1, 16), (2, 8), (4, 8): 4, 2, 1
}

// Largest group size of the shape no larger than min(left, gmax).
template <typename TT>
static int next_group(Shape sh, int left, int gmax) {
  for (const int *g = group_sizes<TT>(sh); *g; ++g)
    if (*g <= left && *g <= gmax) return *g;
  return 1;
}

// Calls F.template run<G, RPW, NW>() for the runtime (g, shape).
template <typename TT, typename F>
static int with_shape(int g, Shape sh, F &&f) {
  if constexpr (sizeof(TT) == 4) {
    if (sh.rpw == 1) {
      switch (g) {
        case 10: return f.template run<10, 1, 16>();
        case 8: return f.template run<8, 1, 16>();
        case 6: return f.template run<6, 1, 16>();
        case 4: return f.template run<4, 1, 16>();
        case 2: return f.template run<2, 1, 16>();
        default: return f.template run<1, 1, 16>();
      }
    }
    if (sh.nw == 12) {
      switch (g) {
        case 5: return f.template run<5, 2, 12>();
        case 2: return f.template run<2, 2, 12>();
        default: return f.template run<1, 2, 12>();
      }
    }
    if (sh.nw == 16) {
      switch (g) {
        case 4: return f.template run<4, 2, 16>();
        case 2: return f.template run<2, 2, 16>();
        default: return f.template run<1, 2, 16>();
      }
    }
  } else {
    if (sh.nw == 12) {
      switch (g) {
        case 2: return f.template run<2, 2, 12>();
        default: return f.template run<1, 2, 12>();
      }
    }
    if (sh.rpw == 1) {
      switch (g) {
        case 4: return f.template run<4, 1, 16>();
        case 2: return f.template run<2, 1, 16>();
        default: return f.template run<1, 1, 16>();
      }
    }
  }
  if (sh.rpw == 4) {
    switch (g) {
      case 4: return f.template run<4, 4, 8>();
      case 2: return f.template run<2, 4, 8>();
      default: return f.template run<1, 4, 8>();
    }
  }
  switch (g) {
    case 4: return f.template run<4, 2, 8>();
    case 2: return f.template run<2, 2, 8>();
    default: return f.template run<1, 2, 8>();
  }
}

// FP32 tracers with other th_tendency contributions: the last operation of
// k_step's FP32 update, out = out + (float)th * (float)(dtime / rho(taup1)),
// over every level of the compute domain, as a second pass.
__global__ void k_add_th_f32(KDom d, float *__restrict__ out, const double *__restrict__ th,
                             const double *__restrict__ R, const double *__restrict__ Rr,
                             double dtime) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.ni) return;
  const int o = d2(d, i, blockIdx.y) + d.P2 * blockIdx.z;
  const double rr = Rr ? Rr[o] : 1.0 / (R[o] + EPSLN);
  const float c32 = (float)(dtime * rr);
  out[o] = out[o] + (float)th[o] * c32;
}

// (launch errors surface at the caller's next launch check)
static void add_th_f32(hipStream_t s, const KDom &d, float *out, const double *th,
                       const double *R, const double *Rr, double dtime) {
  hipLaunchKernelGGL(k_add_th_f32, dim3((d.ni + 255) / 256, d.nj, d.nk), dim3(256), 0, s, d, out,
                     th, R, Rr, dtime);
}

template <typename TT>
struct GroupLaunch {
  mom_gpu_ctx *ctx;
  const StepTab<TT> &tab;
  const double *u, *v, *w, *rho, *rm1, *rr;
  double dtime;
  template <int G, int RPW, int NW>
  int run() const {
    constexpr int TY = NW * RPW - 4;
    const KDom &d = ctx->kd;
    const int ntx = (d.ni + TX - 1) / TX;
    const int nty = (d.nj + TY - 1) / TY;
    const int ntiles = ntx * nty;
    const int per = (ntiles + 7) / 8;
    const StepJob<TT> J{d, tab, u, v, w, rho, rm1, rr, ctx->kmt, ctx->dat, ctx->datr,
                        ctx->dxte, ctx->dyte, ctx->dxtn, ctx->dytn, dtime, ntx, ntiles};
    hipLaunchKernelGGL((k_step<TT, G, RPW, NW>), dim3(8 * per), dim3(RX * NW), 0, ctx->stream,
                       J, per);
    return MOM_GPU_OK;
  }
};

template <typename TT>
static void launch_all(mom_gpu_ctx *ctx, int gmax, int rpw, int waves, int ntr,
                       const void *const *T_in, void *const *T_out, const int *dtype,
                       const double *const *th_in, int want, const double *u, const double *v,
                       const double *w, const double *rho, const double *rm1, const double *rr,
                       double dtime) {
  const Shape sh = shape_of<TT>(rpw, waves);
  std::vector<int> idx;
  for (int n = 0; n < ntr; ++n)
    if (dtype[n] == want) idx.push_back(n);
  size_t i = 0;
  while (i < idx.size()) {
    const int c = next_group<TT>(sh, (int)(idx.size() - i), gmax);
    StepTab<TT> t{};
    for (int q = 0; q < c; ++q) {
      const int n = idx[i + q];
      t.T[q] = (const TT *)T_in[n];
      t.out[q] = (TT *)T_out[n];
      t.th[q] = (th_in && sizeof(TT) == 8) ? th_in[n] : nullptr;
    }
    with_shape<TT>(c, sh, GroupLaunch<TT>{ctx, t, u, v, w, rho, rm1, rr, dtime});
    if (sizeof(TT) == 4 && th_in)
      for (int q = 0; q < c; ++q)
        if (th_in[idx[i + q]])
          add_th_f32(ctx->stream, ctx->kd, (float *)T_out[idx[i + q]], th_in[idx[i + q]], rho,
                     rr, dtime);
    i += c;
  }
}


// ---- ensembles: one queue launch per (storage type, group size) -------------
template <typename TT, int G, int RPW, int NW>
static int queue_launch(mom_gpu_ensemble *e, std::vector<StepJob<TT>> &jobs, int slot) {
  if (jobs.empty()) return MOM_GPU_OK;
  // largest first; the table is compared byte for byte, so jobs are copied
  // as bytes from their zero-initialised slots (no stray padding)
  const size_t nj = jobs.size();
  std::vector<size_t> order(nj);
  for (size_t j = 0; j < nj; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return jobs[a].ntiles > jobs[b].ntiles; });
  const size_t jbytes = (sizeof(StepJob<TT>) * nj + 255) / 256 * 256;
  std::vector<unsigned char> buf(jbytes + sizeof(int) * nj, 0);
  int *first = (int *)(buf.data() + jbytes);
  int total = 0;
  for (size_t j = 0; j < nj; ++j) {
    std::memcpy(buf.data() + sizeof(StepJob<TT>) * j, &jobs[order[j]], sizeof(StepJob<TT>));
    first[j] = total;
    total += jobs[order[j]].ntiles;
  }
  const unsigned char *dev = (const unsigned char *)ens_table(e, buf.data(), buf.size());
  if (!dev) return MOM_GPU_ERR_HIP;
  // enough persistent groups to fill the device; late ones find the queue empty
  const int grid = std::min(total, 2048);
  hipLaunchKernelGGL((k_step_queue<TT, G, RPW, NW>), dim3(grid), dim3(RX * NW), 0,
                     e->lead->stream, (const StepJob<TT> *)dev, (const int *)(dev + jbytes),
                     (int)nj, total, e->counters + slot);
  MOM_LAUNCHED(e->lead, "k_step_queue");
  return MOM_GPU_OK;
}

template <typename TT>
struct QueueLaunch {
  mom_gpu_ensemble *e;
  std::vector<StepJob<TT>> &jobs;
  int slot;
  template <int G, int RPW, int NW>
  int run() const { return queue_launch<TT, G, RPW, NW>(e, jobs, slot); }
};

template <typename TT>
static int ensemble_type(mom_gpu_ensemble *e, const mom_gpu_tracer_step_t *steps, int want,
                         int gmax, int rpw, int waves, int &slot) {
  // jobs bucketed by group size (the shape's sizes no larger than gmax), per domain
  const Shape sh = shape_of<TT>(rpw, waves);
  constexpr int GMAX = 12;
  std::vector<StepJob<TT>> by[GMAX + 1];
  const int TY = sh.nw * sh.rpw - 4;
  for (size_t dd = 0; dd < e->ctx.size(); ++dd) {
    mom_gpu_ctx *ctx = e->ctx[dd];
    const mom_gpu_tracer_step_t &s = steps[dd];
    const KDom &d = ctx->kd;
    const double *rm1 = s.rho_dzt_taum1 ? s.rho_dzt_taum1 : s.rho_dzt_tau;
    std::vector<int> idx;
    for (int n = 0; n < s.ntr; ++n)
      if (s.dtype[n] == want) idx.push_back(n);
    size_t i = 0;
    while (i < idx.size()) {
      const int c = next_group<TT>(sh, (int)(idx.size() - i), gmax);
      StepTab<TT> t{};
      for (int q = 0; q < c; ++q) {
        const int n = idx[i + q];
        t.T[q] = (const TT *)s.T_taum1[n];
        t.out[q] = (TT *)s.T_taup1[n];
        t.th[q] = (s.th_tendency && sizeof(TT) == 8) ? s.th_tendency[n] : nullptr;
      }
      const int ntx = (d.ni + TX - 1) / TX;
      const int ntiles = ntx * ((d.nj + TY - 1) / TY);
      by[c].emplace_back();
      StepJob<TT> &J = by[c].back();
      std::memset((void *)&J, 0, sizeof J);
      J.d = d;
      J.tab = t;
      J.U = s.uhrho_et; J.V = s.vhrho_nt; J.W = s.wrho_bt; J.R = s.rho_dzt_tau; J.Rm1 = rm1;
      J.Rr = s.rho_dztr;
      J.kmt = ctx->kmt; J.dat = ctx->dat; J.datr = ctx->datr; J.dxte = ctx->dxte;
      J.dyte = ctx->dyte; J.dxtn = ctx->dxtn; J.dytn = ctx->dytn;
      J.dtime = s.dtime; J.ntx = ntx; J.ntiles = ntiles;
      i += c;
    }
  }
  int rc = MOM_GPU_OK;
  for (const int *g = group_sizes<TT>(sh); *g && rc == MOM_GPU_OK; ++g)
    if (!by[*g].empty()) rc = with_shape<TT>(*g, sh, QueueLaunch<TT>{e, by[*g], slot++});
  if (rc || sizeof(TT) == 8) return rc;
  bool added = false;
  for (size_t dd = 0; dd < e->ctx.size(); ++dd) {
    const mom_gpu_tracer_step_t &s = steps[dd];
    if (!s.th_tendency) continue;
    for (int n = 0; n < s.ntr; ++n)
      if (s.dtype[n] == want && s.th_tendency[n]) {
        add_th_f32(e->lead->stream, e->ctx[dd]->kd, (float *)s.T_taup1[n], s.th_tendency[n],
                   s.rho_dzt_tau, s.rho_dztr, s.dtime);
        added = true;
      }
  }
  if (added) MOM_LAUNCHED(e->lead, "k_add_th_f32");
  return rc;
}
}  // namespace

// Advection + explicit update for ntr MDFL-Sweby tracers.  dtype[n]: 8 = FP64
// (reference rounding), 4 = FP32 (single precision, division-free limiter).
int launch_tracer_step(mom_gpu_ctx *ctx, int ntr, const void *const *T_in, void *const *T_out,
                       const int *dtype, const double *const *th_in, const double *u,
                       const double *v, const double *w, const double *rho,
                       const double *rho_taum1, const double *rho_dztr, double dtime) {
  launch_all<double>(ctx, ctx->tune.step_group_f64, ctx->tune.step_rows_per_wave,
                     ctx->tune.step_waves_f64, ntr, T_in, T_out, dtype, th_in, 8, u, v, w, rho,
                     rho_taum1, rho_dztr, dtime);
  launch_all<float>(ctx, ctx->tune.step_group_f32, ctx->tune.step_rows_per_wave_f32,
                    ctx->tune.step_waves_f32, ntr, T_in, T_out, dtype, th_in, 4, u, v, w, rho,
                    rho_taum1, rho_dztr, dtime);
  MOM_LAUNCHED(ctx, "k_step");
  return MOM_GPU_OK;
}

// The tracer passes of every domain: one persistent queue launch per storage
// type and group size (knobs from the lead context).
int launch_step_ensemble(mom_gpu_ensemble *e, const mom_gpu_tracer_step_t *steps) {
  mom_gpu_ctx *L = e->lead;
  int rc = launch_set_ints(L, e->counters, 0, mom_gpu_ensemble::NCOUNT);
  if (rc) return rc;
  int slot = 0;
  rc = ensemble_type<double>(e, steps, 8, L->tune.step_group_f64, L->tune.step_rows_per_wave,
                             L->tune.step_waves_f64, slot);
  if (rc) return rc;
  return ensemble_type<float>(e, steps, 4, L->tune.step_group_f32, L->tune.step_rows_per_wave_f32,
                              L->tune.step_waves_f32, slot);
}
