// Column and point-wise kernels: invtri and the tracer loop of
// vert_diffuse_implicit, compute_tmask_limit, the explicit tracer update,
// the masked tracer content and the grid mask check.
#include "mom_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr double EPS_TRI = 1.0e-30;   // ocean_util.F90:175

// ---- invtri, ocean_util.F90:153-222 -------------------------------------
// One thread per column, i fastest across the wavefront so every level is
// one coalesced row.  The forward sweep writes z(k) in place and e(k) to a
// work array; the back substitution re-reads both.
__global__ void k_invtri(KDom d, double *__restrict__ z, const double *__restrict__ topbc,
                         const double *__restrict__ botbc, const double *__restrict__ dcb,
                         double tdt, const int *__restrict__ kmz,
                         const double *__restrict__ mask, const double *__restrict__ dh,
                         const double *__restrict__ dhw, double aidif,
                         double *__restrict__ ework) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  const int gj = blockIdx.y;
  if (gi >= d.ni) return;
  const int nk = d.nk, P2 = d.P2;
  const int o = d2(d, gi, gj);
  const double tdt_aidif = tdt * aidif;
  const int kb = max(2, kmz[o]);
  double cprev = 0.0, bet = 0.0, zprev = 0.0;
  for (int k = 1; k <= nk; ++k) {
    const int km1 = max(1, k - 1), kp1 = min(k + 1, nk);
    const int ok = o + P2 * (k - 1);
    const double dhk = dh[ok];
    const double factu = tdt_aidif / (dhw[o + P2 * (k - 1)] * dhk);
    const double factl = tdt_aidif / (dhw[o + P2 * k] * dhk);
    const double mkk = mask[ok];
    double a = -dcb[o + P2 * (km1 - 1)] * factu * mkk;
    double c = -dcb[ok] * factl * mask[o + P2 * (kp1 - 1)];
    if (k == 1) a = 0.0;
    if (k == nk) c = 0.0;
    const double b = 1.0 - a - c;
    double f;
    if (k == 1)
      f = z[ok] + topbc[o] * tdt_aidif * mkk / dhk;
    else if (k == kb)
      f = z[ok] - botbc[o] * tdt_aidif * mkk / dhk;
    else
      f = z[ok] * mkk;
    double zk;
    if (k == 1) {
      bet = mkk / (b + EPS_TRI);
      zk = f * bet;
    } else {
      const double e = cprev * bet;
      bet = mkk / (b - a * e + EPS_TRI);
      zk = (f - a * zprev) * bet;
      ework[ok] = e;
    }
    z[ok] = zk;
    zprev = zk;
    cprev = c;
  }
  double znext = zprev;
  for (int k = nk - 1; k >= 1; --k) {
    const int ok = o + P2 * (k - 1);
    const double zk = z[ok] - ework[ok + P2] * znext;
    z[ok] = zk;
    znext = zk;
  }
}

// ---- vert_diffuse_implicit, ocean_vert_mix.F90:3462-3491 -----------------
// dcb = rho0*diff_cbt(nmix) + K33 formed on the fly; up to NT tracers that
// share the same dcb share one factorisation (e, bet), identical arithmetic
// to a per-tracer invtri.  The arithmetic is FP64 for every tracer; an FP32
// tracer (TT = float) is read and written in single precision, including
// the forward-sweep values that the back substitution re-reads.
constexpr int NT = 8;
template <typename TT>
struct TriGroup {
  TT *z[NT];
  const double *stf[NT];
  const double *btf[NT];
  const double *k33;   // NULL: zero (only for a group of one tracer)
  int n;
};

// One group of one domain (the ensemble launches read these from a table).
template <typename TT>
struct VdJob {
  KDom d;
  TriGroup<TT> g;
  const double *dcbase;
  double rho0;
  const int *kmt;
  const double *dh, *dhw;
  double tdt, aidif;
  double *ework;
  int nbx, nblocks;
  // k_vdiff_twin64 only: tracers n >= n2 of the group take their diffusivity
  // from dcbase2 (a job holding both mixing classes; other kernels: n2 = 0
  // and dcbase2 unused)
  const double *dcbase2;
  int n2;
};

template <typename TT>
__device__ __forceinline__ void vdiff_col(const VdJob<TT> &J, int gi, int gj) {
  const KDom d = J.d;
  const TriGroup<TT> &g = J.g;
  const double *__restrict__ dcbase = J.dcbase;
  const double rho0 = J.rho0;
  const int *__restrict__ kmt = J.kmt;
  const double *__restrict__ dh = J.dh;
  const double *__restrict__ dhw = J.dhw;
  const double tdt = J.tdt, aidif = J.aidif;
  double *__restrict__ ework = J.ework;
  const int nk = d.nk, P2 = d.P2;
  const int o = d2(d, gi, gj);
  const double tdt_aidif = tdt * aidif;
  const int kmo = kmt[o];
  const int kb = max(2, kmo);
  double cprev = 0.0, bet = 0.0;
  double zprev[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) zprev[n] = 0.0;
  double dcb_km1 = 0.0;
  for (int k = 1; k <= nk; ++k) {
    const int kp1 = min(k + 1, nk);
    const int ok = o + P2 * (k - 1);
    const double dcb_k = rho0 * dcbase[ok] + (g.k33 ? g.k33[ok] : 0.0);
    if (k == 1) dcb_km1 = dcb_k;
    const double dhk = dh[ok];
    const double factu = tdt_aidif / (dhw[o + P2 * (k - 1)] * dhk);
    const double factl = tdt_aidif / (dhw[o + P2 * k] * dhk);
    const double mkk = (k <= kmo) ? 1.0 : 0.0;
    const double mkp1 = (kp1 <= kmo) ? 1.0 : 0.0;
    double a = -dcb_km1 * factu * mkk;
    double c = -dcb_k * factl * mkp1;
    if (k == 1) a = 0.0;
    if (k == nk) c = 0.0;
    const double b = 1.0 - a - c;
    double e = 0.0;
    if (k == 1) {
      bet = mkk / (b + EPS_TRI);
    } else {
      e = cprev * bet;
      bet = mkk / (b - a * e + EPS_TRI);
      ework[ok] = e;
    }
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      if (n < g.n) {
        TT *z = g.z[n];
        const double zo = (double)z[ok];
        double f;
        if (k == 1)
          f = zo + g.stf[n][o] * tdt_aidif * mkk / dhk;
        else if (k == kb)
          f = zo - g.btf[n][o] * tdt_aidif * mkk / dhk;
        else
          f = zo * mkk;
        const double zk = (k == 1) ? f * bet : (f - a * zprev[n]) * bet;
        z[ok] = (TT)zk;
        zprev[n] = zk;
      }
    }
    cprev = c;
    dcb_km1 = dcb_k;
  }
  for (int k = nk - 1; k >= 1; --k) {
    const int ok = o + P2 * (k - 1);
    const double e1 = ework[ok + P2];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      if (n < g.n) {
        TT *z = g.z[n];
        const double zk = (double)z[ok] - e1 * zprev[n];
        z[ok] = (TT)zk;
        zprev[n] = zk;
      }
    }
  }
}

// vdiff_col for a group of one tracer (TT = double), the deepest TAIL levels
// kept in registers between the sweeps: the forward sweep reaches them last
// and the back substitution needs them first, so their z and e never go to
// memory (the same arithmetic, in the same order; nk > TAIL).
template <int TAIL>
__device__ __forceinline__ void vdiff_col1(const VdJob<double> &J, int gi, int gj) {
  const KDom d = J.d;
  const TriGroup<double> &g = J.g;
  const double *__restrict__ dcbase = J.dcbase;
  const double rho0 = J.rho0;
  const double *__restrict__ dh = J.dh;
  const double *__restrict__ dhw = J.dhw;
  const double tdt_aidif = J.tdt * J.aidif;
  double *__restrict__ ework = J.ework;
  double *__restrict__ z = g.z[0];
  const int nk = d.nk, P2 = d.P2;
  const int o = d2(d, gi, gj);
  const int kmo = J.kmt[o];
  const int kb = max(2, kmo);
  const double stf = g.stf[0][o], btf = g.btf[0][o];
  double cprev = 0.0, bet = 0.0, zprev = 0.0, dcb_km1 = 0.0;
  // one level of the forward sweep (vdiff_col's expressions); returns e(k)
  auto level = [&](int k, double &zk) -> double {
    const int kp1 = min(k + 1, nk);
    int ok = o + P2 * (k - 1);
    // the level's addresses wait for the previous level's result, so that
    // the unrolled tail issues one level's loads at a time
    asm volatile("" : "+v"(ok) : "v"(zprev));
    const double dcb_k = rho0 * dcbase[ok] + (g.k33 ? g.k33[ok] : 0.0);
    if (k == 1) dcb_km1 = dcb_k;
    const double dhk = dh[ok];
    const double factu = tdt_aidif / (dhw[ok] * dhk);
    const double factl = tdt_aidif / (dhw[ok + P2] * dhk);
    const double mkk = (k <= kmo) ? 1.0 : 0.0;
    const double mkp1 = (kp1 <= kmo) ? 1.0 : 0.0;
    double a = -dcb_km1 * factu * mkk;
    double c = -dcb_k * factl * mkp1;
    if (k == 1) a = 0.0;
    if (k == nk) c = 0.0;
    const double b = 1.0 - a - c;
    double e = 0.0;
    if (k == 1) {
      bet = mkk / (b + EPS_TRI);
    } else {
      e = cprev * bet;
      bet = mkk / (b - a * e + EPS_TRI);
    }
    const double zo = z[ok];
    double f;
    if (k == 1)
      f = zo + stf * tdt_aidif * mkk / dhk;
    else if (k == kb)
      f = zo - btf * tdt_aidif * mkk / dhk;
    else
      f = zo * mkk;
    zk = (k == 1) ? f * bet : (f - a * zprev) * bet;
    zprev = zk;
    cprev = c;
    dcb_km1 = dcb_k;
    return e;
  };
  const int kt = nk - TAIL;
  for (int k = 1; k <= kt; ++k) {
    double zk;
    const double e = level(k, zk);
    const int ok = o + P2 * (k - 1);
    if (k > 1) ework[ok] = e;
    z[ok] = zk;
  }
  double et[TAIL], zt[TAIL];
#pragma unroll
  for (int t = 0; t < TAIL; ++t) et[t] = level(kt + 1 + t, zt[t]);
  z[o + P2 * (nk - 1)] = zt[TAIL - 1];
  zprev = zt[TAIL - 1];
#pragma unroll
  for (int t = TAIL - 2; t >= 0; --t) {
    const double zk = zt[t] - et[t + 1] * zprev;
    z[o + P2 * (kt + t)] = zk;
    zprev = zk;
  }
  for (int k = kt; k >= 1; --k) {
    const int ok = o + P2 * (k - 1);
    const double e1 = (k == kt) ? et[0] : ework[ok + P2];
    const double zk = z[ok] - e1 * zprev;
    z[ok] = zk;
    zprev = zk;
  }
}

// Register tails: 40 levels (about 210 VGPRs, two waves a SIMD) for nk > 40,
// 20 for 20 < nk <= 40; shallower columns take k_vdiff.  At 3600x2700x75 the
// 40-level tail takes one FP64 tracer from 9.9 to 8.4 ms (the plane traffic of
// 35 levels instead of 55).
constexpr int VD_TAIL = 20;
constexpr int VD_TAIL_DEEP = 40;

template <int TAIL>
__global__ __launch_bounds__(64) void k_vdiff1(VdJob<double> J, int per_xcd) {
  const int b = blockIdx.x;
  const int tile = (b & 7) * per_xcd + (b >> 3);
  if (tile >= J.nblocks) return;
  const int gi = (tile % J.nbx) * 64 + threadIdx.x;
  if (gi >= J.d.ni) return;
  vdiff_col1<TAIL>(J, gi, tile / J.nbx);
}

// 64-column tiles in XCD-aware order (see k_vdiff_twin): J.nbx tiles a row,
// J.nblocks in all, per_xcd consecutive tiles per XCD.
template <typename TT>
__global__ void k_vdiff(VdJob<TT> J, int per_xcd) {
  const int b = blockIdx.x;
  const int tile = (b & 7) * per_xcd + (b >> 3);
  if (tile >= J.nblocks) return;
  const int gi = (tile % J.nbx) * 64 + threadIdx.x;
  if (gi >= J.d.ni) return;
  vdiff_col(J, gi, tile / J.nbx);
}

// FP32 storage: the same tridiagonal system solved for the increment,
//   M dT = bflux - (M - I) T,   T(new) = (T + dT) * tmask,
// with (M - I) T = a (T(k-1) - T(k)) + c (T(k+1) - T(k)) in FP64.  The forward
// values dT' are small numbers kept in an FP32 scratch field, not near-copies
// of T written over the tracer: rounding those (weak-diffusion rows of M are
// close to the identity) is biased and drifts the tracer content, while the
// rounding of dT' is not.  Arithmetic stays FP64.
constexpr int NTF = 12;     // FP32 tracers per implicit-diffusion pass
constexpr int NTF_ENS = 4;  // the same in the ensemble kernel (job tables cost registers)
struct TriGroupF {
  float *z[NTF];
  const double *stf[NTF];
  const double *btf[NTF];
  int n;
};

struct VdJobF {
  KDom d;
  TriGroupF g;
  const double *dcbase;
  double rho0;
  const int *kmt;
  const double *dh, *dhw;
  double tdt, aidif;
  double *ework;
  float *scr;
  int nbx, nblocks;
};

// PF = true: the loads of level k+1 (k-1 on the way back) are issued before
// level k is computed, so that each column has two levels of loads in
// flight (same arithmetic, same results).
template <int MAXN, bool PF>
__device__ __forceinline__ void vdiff_inc_col(const VdJobF &J, int gi, int gj) {
  const KDom d = J.d;
  const TriGroupF &g = J.g;
  const double *__restrict__ dcbase = J.dcbase;
  const double rho0 = J.rho0;
  const int *__restrict__ kmt = J.kmt;
  const double *__restrict__ dh = J.dh;
  const double *__restrict__ dhw = J.dhw;
  const double tdt = J.tdt, aidif = J.aidif;
  double *__restrict__ ework = J.ework;
  float *__restrict__ scr = J.scr;
  const int nk = d.nk, P2 = d.P2;
  const size_t P3 = (size_t)P2 * nk;
  const int o = d2(d, gi, gj);
  const double tdt_aidif = tdt * aidif;
  const int kmo = kmt[o];
  const int kb = max(2, kmo);
  double cprev = 0.0, bet = 0.0, dcb_km1 = 0.0;
  double dprev[MAXN];
  float zkm1[MAXN], zk[MAXN], zn[MAXN];   // FP32 storage values (exact in FP64)
#pragma unroll
  for (int n = 0; n < MAXN; ++n) {
    dprev[n] = 0.0;
    zkm1[n] = 0.0f;
    zk[n] = (n < g.n) ? g.z[n][o] : 0.0f;
    zn[n] = (PF && n < g.n && nk > 1) ? g.z[n][o + P2] : 0.0f;
  }
  // level-k inputs (PF: loaded one level ahead)
  double dcb_n = dcbase[o], dh_n = dh[o], dhwu_n = dhw[o], dhwl_n = dhw[o + P2];
  for (int k = 1; k <= nk; ++k) {
    const int kp1 = min(k + 1, nk);
    const int ok = o + P2 * (k - 1);
    double dcbk, dhk, dhwu, dhwl;
    float zkp1v[MAXN];
    if constexpr (PF) {
      dcbk = dcb_n; dhk = dh_n; dhwu = dhwu_n; dhwl = dhwl_n;
#pragma unroll
      for (int n = 0; n < MAXN; ++n) zkp1v[n] = zn[n];
      if (k < nk) {
        dcb_n = dcbase[ok + P2]; dh_n = dh[ok + P2];
        dhwu_n = dhwl; dhwl_n = dhw[o + P2 * (k + 1)];
#pragma unroll
        for (int n = 0; n < MAXN; ++n)
          zn[n] = (n < g.n && k + 1 < nk) ? g.z[n][ok + 2 * P2] : 0.0f;
      }
    } else {
      dcbk = dcbase[ok]; dhk = dh[ok];
      dhwu = dhw[o + P2 * (k - 1)]; dhwl = dhw[o + P2 * k];
#pragma unroll
      for (int n = 0; n < MAXN; ++n) zkp1v[n] = (n < g.n && k < nk) ? g.z[n][ok + P2] : 0.0f;
    }
    const double dcb_k = rho0 * dcbk;
    if (k == 1) dcb_km1 = dcb_k;
    const double factu = tdt_aidif / (dhwu * dhk);
    const double factl = tdt_aidif / (dhwl * dhk);
    const double mkk = (k <= kmo) ? 1.0 : 0.0;
    const double mkp1 = (kp1 <= kmo) ? 1.0 : 0.0;
    double a = -dcb_km1 * factu * mkk;
    double c = -dcb_k * factl * mkp1;
    if (k == 1) a = 0.0;
    if (k == nk) c = 0.0;
    const double b = 1.0 - a - c;
    if (k == 1) {
      bet = mkk / (b + EPS_TRI);
    } else {
      const double e = cprev * bet;
      bet = mkk / (b - a * e + EPS_TRI);
      ework[ok] = e;
    }
#pragma unroll
    for (int n = 0; n < MAXN; ++n) {
      if (n < g.n) {
        const float zkp1 = (k < nk) ? zkp1v[n] : 0.0f;
        double bfl = 0.0;
        if (k == 1)
          bfl = g.stf[n][o] * tdt_aidif * mkk / dhk;
        else if (k == kb)
          bfl = -(g.btf[n][o] * tdt_aidif * mkk / dhk);
        const double zc = zk[n];
        const double rhs = bfl - (a * ((double)zkm1[n] - zc) + c * ((double)zkp1 - zc));
        const double dk = (k == 1) ? rhs * bet : (rhs - a * dprev[n]) * bet;
        scr[(size_t)n * P3 + ok] = (float)dk;
        dprev[n] = dk;
        zkm1[n] = zk[n];
        zk[n] = zkp1;
      }
    }
    cprev = c;
    dcb_km1 = dcb_k;
  }
  {
    const int ok = o + P2 * (nk - 1);
    const double mk = (nk <= kmo) ? 1.0 : 0.0;
#pragma unroll
    for (int n = 0; n < MAXN; ++n)
      if (n < g.n) g.z[n][ok] = (float)(((double)g.z[n][ok] + dprev[n]) * mk);
  }
  // back substitution, k = nk-1 .. 1 (PF: level k-1 loaded while k is done)
  float sn[MAXN], tnx[MAXN];
  double en = 0.0;
  if constexpr (PF) {
    if (nk >= 2) {
      const int ok = o + P2 * (nk - 2);
      en = ework[ok + P2];
#pragma unroll
      for (int n = 0; n < MAXN; ++n) {
        sn[n] = (n < g.n) ? scr[(size_t)n * P3 + ok] : 0.0f;
        tnx[n] = (n < g.n) ? g.z[n][ok] : 0.0f;
      }
    }
  }
  for (int k = nk - 1; k >= 1; --k) {
    const int ok = o + P2 * (k - 1);
    double e1;
    float sv[MAXN], tv[MAXN];
    if constexpr (PF) {
      e1 = en;
#pragma unroll
      for (int n = 0; n < MAXN; ++n) { sv[n] = sn[n]; tv[n] = tnx[n]; }
      if (k > 1) {
        const int om = ok - P2;
        en = ework[om + P2];
#pragma unroll
        for (int n = 0; n < MAXN; ++n) {
          sn[n] = (n < g.n) ? scr[(size_t)n * P3 + om] : 0.0f;
          tnx[n] = (n < g.n) ? g.z[n][om] : 0.0f;
        }
      }
    } else {
      e1 = ework[ok + P2];
#pragma unroll
      for (int n = 0; n < MAXN; ++n) {
        sv[n] = (n < g.n) ? scr[(size_t)n * P3 + ok] : 0.0f;
        tv[n] = (n < g.n) ? g.z[n][ok] : 0.0f;
      }
    }
    const double mk = (k <= kmo) ? 1.0 : 0.0;
#pragma unroll
    for (int n = 0; n < MAXN; ++n) {
      if (n < g.n) {
        const double dT = (double)sv[n] - e1 * dprev[n];
        g.z[n][ok] = (float)(((double)tv[n] + dT) * mk);
        dprev[n] = dT;
      }
    }
  }
}

// Launched with 64 or 256 threads a block (knob vdiff_block): four
// wavefronts of one block sweep 256 neighbouring columns together, so each
// level's loads and stores reach HBM as 1-2 KB runs instead of 256-512 B.
template <bool PF>
__global__ __launch_bounds__(256) void k_vdiff_inc(VdJobF J) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= J.d.ni) return;
  vdiff_inc_col<NTF, PF>(J, gi, blockIdx.y);
}

// FP32 tracers, every level of a column in registers: two wavefronts per 64
// columns.  The block's first wavefront holds the top nt = ceil(nk/2) levels
// of each column and eliminates downward (Thomas); the second holds the other
// nb = nk - nt levels and eliminates upward, which is the same recurrence with
// the sub- and super-diagonal swapped.  The two partial eliminations meet at
// the interface between levels nt and nt + 1, where
//     x_top = d_top - E_top x_bot,   x_bot = d_bot - E_bot x_top
// is solved with the other wavefront's (d, E) of the column (through LDS);
// each wavefront then back-substitutes over its own levels.
//
// A wavefront has H register slots; its nact levels fill slots ph..H-1
// (ph = H - nact), so both meet in slot H-1: slot ph + r is level r + 1
// (top) or nk - r (bottom), the same plane for all lanes (a scalar base and
// one lane offset), and slot H holds the other half's last level.  Slots
// below ph are inert rows (P = E = 0, no loads or stores) that precede the
// surface or bottom row, whose coupling to them is zero.
//
// The system is invtri's (ocean_util.F90:175-236) in increment form, as in
// vdiff_inc_col: x = T(taup1) - T.  Coefficients and the factorisation are
// formed in FP64 (the same expressions), kept as FP32 multipliers
// P = p bet, E = q bet (p couples a row to the previous slot, q to the next);
// each tracer's sweeps run in FP32.  Every tracer is read once and written
// once; there is no scratch plane.
// Wavefront-uniform base + 32-bit lane byte offset.  The base is pinned in
// scalar registers, so that each access is one global load/store of the
// scalar-base + vector-offset form instead of a per-slot 64-bit vector
// address that the compiler would otherwise form (and keep) for each slot.
#define MOM_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T ld_at(const T *base, unsigned byte_off) {
  asm volatile("" : "+s"(base));
  return *(const MOM_GLOBAL T *)((const MOM_GLOBAL char *)base + byte_off);
}
template <typename T>
__device__ __forceinline__ void st_at(T *base, unsigned byte_off, T v) {
  asm volatile("" : "+s"(base));
  *(MOM_GLOBAL T *)((MOM_GLOBAL char *)base + byte_off) = v;
}
// the LDS exchange of the two wavefronts: a barrier that orders LDS only, so
// that loads in flight (the next tracer) are not waited for
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One 64-column tile (J.nbx tiles a row) of job J, by wavefront pair
// threadIdx.x / 128 of a block of NP pairs; every thread of the block calls it
// (it holds block barriers: one, then one per tracer).  A pair whose tile is
// past the end (live false) runs a valid tile without storing.
template <int H, int NP = 1>
__device__ __forceinline__ void twin_tile(const VdJobF &J, int tile, bool live = true) {
  __shared__ float xch_all[NP][2][2][64];    // [pair][tracer parity][half][column]
  __shared__ double ech_all[NP][2][64];
  const int pr = NP > 1 ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 7) : 0;
  float (*xch)[2][64] = xch_all[pr];
  double (*ech)[64] = ech_all[pr];
  const KDom d = J.d;
  const int half = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 1);
  const bool top = half == 0;
  const int lane = threadIdx.x & 63;
  const int col = (tile % J.nbx) * 64 + lane;
  const bool valid = live && col < d.ni;
  const int nk = d.nk;
  const int nt = (nk + 1) >> 1;
  const int nact = top ? nt : nk - nt;
  const int ph = H - nact;
  const unsigned o = (unsigned)d2(d, valid ? col : d.ni - 1, tile / J.nbx);
  const unsigned ob4 = o * 4u, ob8 = o * 8u;  // lane byte offsets in a plane
  const long sP = (long)d.P2;
  const long step = top ? sP : -sP;           // plane stride from slot to slot
  const long p0 = top ? 0 : (long)(nk - 1) * sP;   // plane of slot ph
  const int kmo = ld_at(J.kmt, ob4);
  const int kb = max(2, kmo);
  const double tda = J.tdt * J.aidif;
  // slot ph + r: dh plane = its level's; the next-side interface (between
  // its level and the next slot's): diff_cbt plane = upper level - 1,
  // dzwt plane = the interface index.  Inert slots read slot ph's planes.
  // The loads of a chunk of CH slots are issued before the previous chunk
  // is computed (the empty asm keeps later chunks from being hoisted).
  constexpr int CH = 8, NCH = (H + CH - 1) / CH;
  const double *dhp = J.dh + p0;
  const double *dcp = J.dcbase + p0 - (top ? 0 : sP);
  const double *dwp = J.dhw + p0 + (top ? sP : 0);
  double Ldh[NCH * CH], Ldc[NCH * CH], Ldw[NCH * CH];
  auto cload = [&](int c) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int j = c * CH + u;
      const long r = max(j - ph, 0);
      Ldh[j] = ld_at(dhp + r * step, ob8);
      Ldc[j] = ld_at(dcp + r * step, ob8);
      Ldw[j] = ld_at(dwp + r * step, ob8);
    }
  };
  float P[H], E[H];
  double Dp = 0.0, Wp = 1.0, mdp = 0.0, Eprev = 0.0;
  // bet and dh of the surface row (top) and of row kb (bottom flux)
  double bet1 = 0.0, dh1 = 1.0, betkb = 0.0, dhkb = 1.0;
  int jkb = -1;
  cload(0);
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    if (c + 1 < NCH) cload(c + 1);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int j = c * CH + u;
      if (j >= H) break;
      const bool real = j >= ph;              // uniform
      const int r = j - ph;
      const int lev = top ? r + 1 : nk - r;
      double dhk = Ldh[j], Wn = Ldw[j];
      // tie this slot's divisions to the previous slot's result: computed
      // one slot at a time, they are not all held in registers at once
      asm volatile("" : "+v"(dhk), "+v"(Wn) : "v"(Eprev));
      const double Dn = J.rho0 * Ldc[j];
      const double mkk = (real && lev <= kmo) ? 1.0 : 0.0;
      // a coefficient is masked by the deeper of its two levels
      const double mdn = real ? (top ? ((lev + 1 <= kmo) ? 1.0 : 0.0) : mkk) : 0.0;
      const double p = (r <= 0) ? 0.0 : -Dp * (tda / (Wp * dhk)) * mdp;
      const double q = -Dn * (tda / (Wn * dhk)) * mdn;
      const double b = 1.0 - p - q;
      const double bet = mkk / (b - p * Eprev + EPS_TRI);
      Eprev = q * bet;
      // the FP32 multipliers are formed here (not where they are used), so
      // that only they, not their FP64 sources, stay in registers
      float pj = (float)(p * bet), ej = (float)Eprev;
      asm volatile("" : "+v"(pj), "+v"(ej));
      P[j] = pj;
      E[j] = ej;
      if (real && lev == 1) { bet1 = bet; dh1 = dhk; }
      const bool atkb = real && lev == kb;
      betkb = atkb ? bet : betkb;
      dhkb = atkb ? dhk : dhkb;
      jkb = atkb ? j : jkb;
      Dp = Dn; Wp = Wn; mdp = mdn;
      // one slot at a time: the divisions of later slots are not hoisted
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // bfl of invtri's rows 1 and kb, times bet (mask: bet is 0 on dry rows)
  const float sstf = (float)(bet1 * tda / dh1);
  const float sbtf = (float)(-(betkb * tda / dhkb));
  ech[half][lane] = Eprev;
  lds_barrier();
  const double Eo = ech[1 - half][lane];
  const float Es = (float)Eprev;
  const float den = (float)(1.0 / (1.0 - Eprev * Eo));
  // wet slots: top j < ph + kmo, bottom j >= ph + nk - kmo
  const int wlo = top ? 0 : ph + nk - kmo, whi = top ? ph + kmo : H;
  const int js = top ? ph : -1;               // the surface row

  float T[H + 1], Tn[H + 1];
  double stf = 0.0, btf = 0.0, stfn = 0.0, btfn = 0.0;
  // branch-free: inert slots read slot ph's level again (their value is
  // not used; P = E = 0 there)
  auto load = [&](int n, int phl, float *dst, double &s, double &bt) {
    const float *zp = J.g.z[n] + p0;
#pragma unroll
    for (int j = 0; j <= H; ++j) {
      dst[j] = ld_at(zp, ob4);
      zp += (j >= phl) ? step : 0;
    }
    s = ld_at(J.g.stf[n], ob8);
    bt = ld_at(J.g.btf[n], ob8);
  };
  if (J.g.n > 0) load(0, ph, T, stf, btf);
#pragma nounroll
  for (int n = 0; n < J.g.n; ++n) {
    // slot bounds and thresholds opaque to the compiler, so that the
    // per-slot compares and addresses are formed in the loop instead of
    // being kept in registers across it
    int phl = ph, jsl = js, jkb_l = jkb, wlo_l = wlo, whi_l = whi;
    asm volatile("" : "+s"(phl), "+s"(jsl), "+v"(jkb_l), "+v"(wlo_l), "+v"(whi_l));
    if (n + 1 < J.g.n) load(n + 1, phl, Tn, stfn, btfn);
    const float inj0 = (float)stf * sstf, injb = (float)btf * sbtf;
    float dv[H];
    float dprev = 0.0f;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const float t1 = (j > 0 ? T[j - 1] - T[j] : 0.0f) + dprev;
      const float t2 = T[j + 1] - T[j];
      float dj = fmaf(-P[j], t1, -(E[j] * t2));
      dj += (j == jsl) ? inj0 : 0.0f;
      dj += (j == jkb_l) ? injb : 0.0f;
      dv[j] = dj;
      dprev = dj;
    }
    xch[n & 1][half][lane] = dprev;
    lds_barrier();
    const float dother = xch[n & 1][1 - half][lane];
    float x = (dprev - Es * dother) * den;
    float *zp = J.g.z[n] + p0 + (long)(nact - 1) * step;
#pragma unroll
    for (int j = H - 1; j >= 0; --j) {
      if (j < H - 1) x = fmaf(-E[j], x, dv[j]);   // slot H-1 keeps the meeting value
      if (j >= phl) {
        const bool wet = j >= wlo_l && j < whi_l;
        if (valid) st_at(zp, ob4, wet ? T[j] + x : 0.0f);
        zp -= step;
      }
    }
#pragma unroll
    for (int j = 0; j <= H; ++j) T[j] = Tn[j];
    stf = stfn;
    btf = btfn;
  }
}

// Blocks are dealt to the 8 XCDs round-robin; each XCD takes a contiguous
// run of 64-column tiles, so the cache lines that two neighbouring tiles
// share (rows are not line-aligned) are fetched into one L2.  A block holds
// two neighbouring tiles (per_xcd counts tile pairs), so that the two halves
// of such a line are also written together, before it leaves L2, rather
// than microseconds apart from two blocks as two partial writes.
template <int H>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_vdiff_twin(VdJobF J, int per_xcd) {
  const int b = blockIdx.x;
  const int t0 = 2 * ((b & 7) * per_xcd + (b >> 3));
  if (t0 >= J.nblocks) return;               // the whole block
  const int tile = t0 + (threadIdx.x >> 7);
  twin_tile<H, 2>(J, min(tile, J.nblocks - 1), tile < J.nblocks);
}

// Every tracer of a mixed step in one launch: FP32 tracers (mixing class 1,
// increment form as k_vdiff_inc) and up to two FP64 tracers of each mixing
// class (the full form of vdiff_col, bit for bit).  dh, dhw and the class-1
// factorisation are read and formed once for all of them; class 2 (salt)
// has its own dcb, factorisation and e plane.
constexpr int NT64 = 2;
struct VdMixJob {
  KDom d;
  TriGroupF g32;                     // class 1, FP32 storage
  double *z1[NT64], *z2[NT64];       // FP64 tracers of class 1 and class 2
  const double *stf1[NT64], *btf1[NT64], *stf2[NT64], *btf2[NT64];
  int n1, n2;
  const double *dcb1, *dcb2;         // diff_cbt planes of the two classes
  double rho0;
  const int *kmt;
  const double *dh, *dhw;
  double tdt, aidif;
  double *ew1, *ew2;
  float *scr;
};

// One class's factorisation step at level k (vdiff_col's arithmetic).
struct TriCoef {
  double a, c, bet, cprev, dcb_km1;
};
__device__ inline void tri_level(TriCoef &t, double dcb_k, double factu, double factl, double mkk,
                                 double mkp1, int k, int nk, double *ew, int ok) {
  if (k == 1) t.dcb_km1 = dcb_k;
  double a = -t.dcb_km1 * factu * mkk;
  double c = -dcb_k * factl * mkp1;
  if (k == 1) a = 0.0;
  if (k == nk) c = 0.0;
  const double b = 1.0 - a - c;
  if (k == 1) {
    t.bet = mkk / (b + EPS_TRI);
  } else {
    const double e = t.cprev * t.bet;
    t.bet = mkk / (b - a * e + EPS_TRI);
    ew[ok] = e;
  }
  t.a = a;
  t.c = c;
}

template <bool PF, int N32>
__global__ __launch_bounds__(256) void k_vdiff_mix(VdMixJob J) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x, gj = blockIdx.y;
  const KDom d = J.d;
  if (gi >= d.ni) return;
  const TriGroupF &g = J.g32;
  const int nk = d.nk, P2 = d.P2;
  const size_t P3 = (size_t)P2 * nk;
  const int o = d2(d, gi, gj);
  const double tdt_aidif = J.tdt * J.aidif;
  const int kmo = J.kmt[o];
  const int kb = max(2, kmo);
  const bool two = J.n2 > 0;
  TriCoef t1{0.0, 0.0, 0.0, 0.0, 0.0}, t2{0.0, 0.0, 0.0, 0.0, 0.0};
  double dprev[N32], zp1[NT64], zp2[NT64];
  float zkm1[N32], zk[N32], zn[N32];
#pragma unroll
  for (int n = 0; n < N32; ++n) {
    dprev[n] = 0.0;
    zkm1[n] = 0.0f;
    zk[n] = (n < g.n) ? g.z[n][o] : 0.0f;
    zn[n] = (PF && n < g.n && nk > 1) ? g.z[n][o + P2] : 0.0f;
  }
#pragma unroll
  for (int n = 0; n < NT64; ++n) { zp1[n] = 0.0; zp2[n] = 0.0; }
  double dcb1_n = J.dcb1[o], dcb2_n = two ? J.dcb2[o] : 0.0, dh_n = J.dh[o];
  double dhwu_n = J.dhw[o], dhwl_n = J.dhw[o + P2];
  for (int k = 1; k <= nk; ++k) {
    const int kp1 = min(k + 1, nk);
    const int ok = o + P2 * (k - 1);
    double dcb1k, dcb2k, dhk, dhwu, dhwl;
    float zkp1v[N32];
    if constexpr (PF) {
      dcb1k = dcb1_n; dcb2k = dcb2_n; dhk = dh_n; dhwu = dhwu_n; dhwl = dhwl_n;
#pragma unroll
      for (int n = 0; n < N32; ++n) zkp1v[n] = zn[n];
      if (k < nk) {
        dcb1_n = J.dcb1[ok + P2];
        if (two) dcb2_n = J.dcb2[ok + P2];
        dh_n = J.dh[ok + P2];
        dhwu_n = dhwl;
        dhwl_n = J.dhw[o + P2 * (k + 1)];
#pragma unroll
        for (int n = 0; n < N32; ++n)
          zn[n] = (n < g.n && k + 1 < nk) ? g.z[n][ok + 2 * P2] : 0.0f;
      }
    } else {
      dcb1k = J.dcb1[ok];
      dcb2k = two ? J.dcb2[ok] : 0.0;
      dhk = J.dh[ok];
      dhwu = J.dhw[o + P2 * (k - 1)];
      dhwl = J.dhw[o + P2 * k];
#pragma unroll
      for (int n = 0; n < N32; ++n) zkp1v[n] = (n < g.n && k < nk) ? g.z[n][ok + P2] : 0.0f;
    }
    const double factu = tdt_aidif / (dhwu * dhk);
    const double factl = tdt_aidif / (dhwl * dhk);
    const double mkk = (k <= kmo) ? 1.0 : 0.0;
    const double mkp1 = (kp1 <= kmo) ? 1.0 : 0.0;
    // dcb = rho0*diff_cbt + K33 with K33 = 0 (ocean_vert_mix.F90:3467-3484)
    const double dcb1_k = J.rho0 * dcb1k + 0.0;
    tri_level(t1, dcb1_k, factu, factl, mkk, mkp1, k, nk, J.ew1, ok);
    double dcb2_k = 0.0;
    if (two) {
      dcb2_k = J.rho0 * dcb2k + 0.0;
      tri_level(t2, dcb2_k, factu, factl, mkk, mkp1, k, nk, J.ew2, ok);
    }
    // FP64 tracers: forward sweep of invtri, stored in place
    auto fwd64 = [&](double *z, const double *stf, const double *btf, const TriCoef &t,
                     double &zprev) {
      const double zo = z[ok];
      double f;
      if (k == 1)
        f = zo + stf[o] * tdt_aidif * mkk / dhk;
      else if (k == kb)
        f = zo - btf[o] * tdt_aidif * mkk / dhk;
      else
        f = zo * mkk;
      const double zkv = (k == 1) ? f * t.bet : (f - t.a * zprev) * t.bet;
      z[ok] = zkv;
      zprev = zkv;
    };
#pragma unroll
    for (int n = 0; n < NT64; ++n) {
      if (n < J.n1) fwd64(J.z1[n], J.stf1[n], J.btf1[n], t1, zp1[n]);
      if (n < J.n2) fwd64(J.z2[n], J.stf2[n], J.btf2[n], t2, zp2[n]);
    }
    // FP32 tracers: increment form (k_vdiff_inc)
#pragma unroll
    for (int n = 0; n < N32; ++n) {
      if (n < g.n) {
        const float zkp1 = (k < nk) ? zkp1v[n] : 0.0f;
        double bfl = 0.0;
        if (k == 1)
          bfl = g.stf[n][o] * tdt_aidif * mkk / dhk;
        else if (k == kb)
          bfl = -(g.btf[n][o] * tdt_aidif * mkk / dhk);
        const double zc = zk[n];
        const double rhs = bfl - (t1.a * ((double)zkm1[n] - zc) + t1.c * ((double)zkp1 - zc));
        const double dk = (k == 1) ? rhs * t1.bet : (rhs - t1.a * dprev[n]) * t1.bet;
        J.scr[(size_t)n * P3 + ok] = (float)dk;
        dprev[n] = dk;
        zkm1[n] = zk[n];
        zk[n] = zkp1;
      }
    }
    t1.cprev = t1.c;
    t1.dcb_km1 = dcb1_k;
    if (two) {
      t2.cprev = t2.c;
      t2.dcb_km1 = dcb2_k;
    }
  }
  {
    const int ok = o + P2 * (nk - 1);
    const double mk = (nk <= kmo) ? 1.0 : 0.0;
#pragma unroll
    for (int n = 0; n < N32; ++n)
      if (n < g.n) g.z[n][ok] = (float)(((double)g.z[n][ok] + dprev[n]) * mk);
  }
  // back substitution
  for (int k = nk - 1; k >= 1; --k) {
    const int ok = o + P2 * (k - 1);
    const double e1 = J.ew1[ok + P2];
    const double e2 = two ? J.ew2[ok + P2] : 0.0;
#pragma unroll
    for (int n = 0; n < NT64; ++n) {
      if (n < J.n1) {
        const double zkv = J.z1[n][ok] - e1 * zp1[n];
        J.z1[n][ok] = zkv;
        zp1[n] = zkv;
      }
      if (n < J.n2) {
        const double zkv = J.z2[n][ok] - e2 * zp2[n];
        J.z2[n][ok] = zkv;
        zp2[n] = zkv;
      }
    }
    const double mk = (k <= kmo) ? 1.0 : 0.0;
#pragma unroll
    for (int n = 0; n < N32; ++n) {
      if (n < g.n) {
        const double dT = (double)J.scr[(size_t)n * P3 + ok] - e1 * dprev[n];
        g.z[n][ok] = (float)(((double)g.z[n][ok] + dT) * mk);
        dprev[n] = dT;
      }
    }
  }
}

// Ensemble: one group of every domain per launch (groups of one domain share
// that domain's solver scratch, so they go in successive launches).
constexpr int VD_MAXJ = 64;
struct VdLaunch {
  int first[VD_MAXJ + 1];
  int njobs;
};

template <typename JOB>
__device__ inline const JOB *vd_pick(const JOB *jobs, const VdLaunch &L, int &gi, int &gj) {
  const int b = blockIdx.x;
  if (b >= L.first[L.njobs]) return nullptr;
  int lo = 0, hi = L.njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.first[mid] <= b) lo = mid; else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);   // uniform: the job table is read by scalar loads
  const JOB &J = jobs[lo];
  const int local = b - L.first[lo];
  gi = (local % J.nbx) * 64 + threadIdx.x;
  gj = local / J.nbx;
  return gi < J.d.ni ? &J : nullptr;
}

__global__ __launch_bounds__(64) void k_vdiff_ens(const VdJob<double> *__restrict__ jobs,
                                                  VdLaunch L) {
  int gi, gj;
  const VdJob<double> *J = vd_pick(jobs, L, gi, gj);
  if (J) vdiff_col(*J, gi, gj);
}

__global__ __launch_bounds__(64) void k_vdiff_inc_ens(const VdJobF *__restrict__ jobs, VdLaunch L) {
  int gi, gj;
  const VdJobF *J = vd_pick(jobs, L, gi, gj);
  if (J) vdiff_inc_col<NTF_ENS, false>(*J, gi, gj);
}

// Ensemble, register-resident FP32 solver: every group of every domain in one
// launch (no scratch is shared); tiles in XCD-aware order over the whole
// table, H the largest slot count the domains need.
template <int H>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
void k_vdiff_twin_ens(const VdJobF *__restrict__ jobs, VdLaunch L, int per_xcd) {
  const int b = blockIdx.x;
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= L.first[L.njobs]) return;         // the whole block
  int lo = 0, hi = L.njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.first[mid] <= t) lo = mid; else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  twin_tile<H>(jobs[lo], t - L.first[lo]);
}

// ---- FP64 tracers, every level of a column in registers ---------------------
// The system of invtri (ocean_util.F90:175-236) for FP64 tracers, solved by two
// wavefronts per 64 columns.  The first eliminates levels 1..nt (nt =
// ceil(nk/2)) downward with the reference's recurrence
//   bet_k = mask_k/(b_k - a_k e_k + eps),  e_k = c_{k-1} bet_{k-1},
//   z'_k = (f_k - a_k z'_{k-1}) bet_k,
// the second levels nk..nt+1 upward, the same recurrence with the sub- and
// super-diagonals swapped (p = a, q = c on top; p = c, q = a below):
//   bet_k = mask_k/(b_k - p_k e_k + eps),  e_k = q_{prev} bet_{prev},
//   v_k = (f_k - p_k v_{prev}) bet_k.
// Each wavefront's last level then satisfies v = x + (q bet) x_other; the two
// rows at the meeting are solved with the other wavefront's (v, q bet) through
// LDS, and each wavefront back-substitutes x = v - (q bet) x_next over its own
// levels.  a, b, c, f and every division are invtri's expressions; only the
// order of elimination differs from its Thomas sweep (the matrix is diagonally
// dominant: the two agree to rounding).  v and q bet of every level stay in
// registers, so each tracer is read and written once and there is no
// factorisation plane.  The levels of a wavefront fill slots ph..H-1 (ph = H -
// its level count), so both meet in slot H-1; slots below ph are inert.  A
// group's tracers share the coefficients' addresses, not their factorisation:
// each tracer repeats it (the coefficients come from L2 the second time).
// Wavefront pair threadIdx.x / 128 of a block of NP pairs; a pair whose
// tile is past the end (live false) runs a valid tile without storing.
template <int H, int NP = 1>
__device__ __forceinline__ void twin64_tile(const VdJob<double> &J, int tile, bool live = true) {
  __shared__ double xv_all[NP][2][2][64], xq_all[NP][2][2][64];   // [pair][parity][half][column]
  const int pr = NP > 1 ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 7) : 0;
  double (*xv)[2][64] = xv_all[pr];
  double (*xq)[2][64] = xq_all[pr];
  const KDom d = J.d;
  const int half = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 1);
  const bool top = half == 0;
  const int lane = threadIdx.x & 63;
  const int col = (tile % J.nbx) * 64 + lane;
  const bool valid = live && col < d.ni;
  const int nk = d.nk;
  const int nt = (nk + 1) >> 1;
  const int nact = top ? nt : nk - nt;
  const int ph = H - nact;
  const long P2 = d.P2;
  const long step = top ? P2 : -P2;            // plane stride from slot to slot
  const long p0 = top ? 0 : (long)(nk - 1) * P2;   // plane of slot ph
  const unsigned o = (unsigned)d2(d, valid ? col : d.ni - 1, tile / J.nbx);
  const unsigned ob8 = o * 8u;                 // lane byte offset in a plane
  const int kmo = ld_at(J.kmt, o * 4u);
  const int kb = max(2, kmo);
  const double tda = J.tdt * J.aidif;
  const double rho0 = J.rho0;
  const double *k33 = J.g.k33;
  for (int n = 0; n < J.g.n; ++n) {
    // the tracer's diffusivity plane: its mixing class's diff_cbt
    const double *dcbase = (J.dcbase2 && n >= J.n2) ? J.dcbase2 : J.dcbase;
    // dcb = rho0*diff_cbt(nmix) + K33 of the plane at offset pl (vdiff_col's
    // expression)
    auto dcb_at = [&](long pl) {
      return rho0 * ld_at(dcbase + pl, ob8) + (k33 ? ld_at(k33 + pl, ob8) : 0.0);
    };
    double *z = J.g.z[n];
    const double stf = ld_at(J.g.stf[n], ob8), btf = ld_at(J.g.btf[n], ob8);
    double vs[H], qs[H];
    // the active levels, slot ph's plane first; one running plane offset.  The
    // offset and ph are opaque to the compiler at every slot, so that the
    // slots' offsets and their active/inert tests are formed one at a time
    // instead of all being kept (as scalar pairs: they would spill)
    long pl = p0;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      vs[j] = 0.0;
      int phj = ph;
      asm volatile("" : "+s"(phj));
      if (j < phj) continue;           // inert slot (wavefront-uniform)
      asm volatile("" : "+s"(pl));
      vs[j] = ld_at(z + pl, ob8);
      pl += step;
    }
    // The coefficients of a chunk of CH slots are loaded while the previous
    // chunk is computed (its divisions wait for them otherwise): per slot the
    // level's dh, the dzwt interface and the diff_cbt plane that the slot
    // adds (top: the interface below and the level's plane; bottom: the
    // interface above and the plane above), with the first slot's other
    // interface and plane loaded before the loop.  Inert slots load slot ph's.
    constexpr int CH = 4, NCH = (H + CH - 1) / CH;
    const long dwo = top ? P2 : 0, dco = top ? 0 : -P2;
    double Ldh[NCH * CH], Ldw[NCH * CH], Ldc[NCH * CH];
    auto cload = [&](int c) {
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int j = c * CH + u;
        if (j >= H) break;
        int phj = ph;
        asm volatile("" : "+s"(phj));
        long plj = p0 + (long)max(j - phj, 0) * step;
        asm volatile("" : "+s"(plj));
        Ldh[j] = ld_at(J.dh + plj, ob8);
        Ldw[j] = ld_at(J.dhw + plj + dwo, ob8);
        Ldc[j] = ld_at(dcbase + plj + dco, ob8);
      }
    };
    // the first slot's other interface and plane (top: dzwt(0) and none; the
    // bottom: dzwt(nk) and diff_cbt(nk))
    const double dhw_first = ld_at(J.dhw + p0 + (top ? 0 : P2), ob8);
    const double dcb_first = top ? 0.0 : dcb_at(p0);
    double bet = 0.0, qprev = 0.0, vprev = 0.0;
    double dcb_c = 0.0, dhw_c = 0.0;   // the interface shared with the next slot
    pl = p0;
    cload(0);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      if (c + 1 < NCH) cload(c + 1);
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int j = c * CH + u;
        if (j >= H) break;
        qs[j] = 0.0;
        int phj = ph;
        asm volatile("" : "+s"(phj));
        if (j < phj) continue;
        asm volatile("" : "+s"(pl));
        const int r = j - phj;
        const int k = top ? r + 1 : nk - r;     // pl = P2*(k - 1)
        const double dhk = Ldh[j];
        // the slot's diff_cbt plane: rho0*diff_cbt (+ K33, read here when a
        // tracer has its own)
        const double dcb_new = J.rho0 * Ldc[j] + (k33 ? ld_at(k33 + pl + dco, ob8) : 0.0);
        // dzwt plane k-1 is the interface above level k, plane k the one below
        double dcb_km1, dcb_k, dhw_km1, dhw_k;
        if (top) {
          dcb_k = dcb_new;
          dhw_k = Ldw[j];
          dcb_km1 = (r == 0) ? dcb_k : dcb_c;
          dhw_km1 = (r == 0) ? dhw_first : dhw_c;
          dcb_c = dcb_k;
          dhw_c = dhw_k;
        } else {
          dcb_k = (r == 0) ? dcb_first : dcb_c;
          dhw_k = (r == 0) ? dhw_first : dhw_c;
          dcb_km1 = dcb_new;
          dhw_km1 = Ldw[j];
          dcb_c = dcb_km1;
          dhw_c = dhw_km1;
        }
        pl += step;
        const int kp1 = min(k + 1, nk);
        const double factu = tda / (dhw_km1 * dhk);
        const double factl = tda / (dhw_k * dhk);
        const double mkk = (k <= kmo) ? 1.0 : 0.0;
        const double mkp1 = (kp1 <= kmo) ? 1.0 : 0.0;
        double a = -dcb_km1 * factu * mkk;
        double c2 = -dcb_k * factl * mkp1;
        if (k == 1) a = 0.0;
        if (k == nk) c2 = 0.0;
        const double b = 1.0 - a - c2;
        const double zo = vs[j];
        double f;
        if (k == 1)
          f = zo + stf * tda * mkk / dhk;
        else if (k == kb)
          f = zo - btf * tda * mkk / dhk;
        else
          f = zo * mkk;
        const double p = top ? a : c2, q = top ? c2 : a;
        double v;
        if (r == 0) {
          bet = mkk / (b + EPS_TRI);
          v = f * bet;
        } else {
          const double e = qprev * bet;
          bet = mkk / (b - p * e + EPS_TRI);
          v = (f - p * vprev) * bet;
        }
        vs[j] = v;
        qs[j] = q * bet;
        qprev = q;
        vprev = v;
        // one slot at a time: later slots' divisions are not hoisted
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // the meeting: x_t + Qt x_b = Vt (top's last level), x_b + Qb x_t = Vb
    xv[n & 1][half][lane] = vs[H - 1];
    xq[n & 1][half][lane] = qs[H - 1];
    lds_barrier();
    const double Vo = xv[n & 1][1 - half][lane], Qo = xq[n & 1][1 - half][lane];
    const double Vt = top ? vs[H - 1] : Vo, Qt = top ? qs[H - 1] : Qo;
    const double Vb = top ? Vo : vs[H - 1], Qb = top ? Qo : qs[H - 1];
    const double xt = (Vt - Qt * Vb) / (1.0 - Qt * Qb);
    double x = top ? xt : Vb - Qb * xt;
    // back substitution from slot H-1 (plane p0 + (nact-1)*step) outward
    pl = p0 + (long)(nact - 1) * step;
#pragma unroll
    for (int j = H - 1; j >= 0; --j) {
      int phj = ph;
      asm volatile("" : "+s"(phj));
      if (j < phj) continue;
      if (j < H - 1) x = vs[j] - qs[j] * x;
      asm volatile("" : "+s"(pl));
      if (valid) st_at(z + pl, ob8, x);
      pl -= step;
    }
  }
}

// Two neighbouring tiles a block (per_xcd counts tile pairs), as for the
// FP32 solver: the halves of the cache lines they share are written together.
template <int H>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_vdiff_twin64(VdJob<double> J, int per_xcd) {
  const int b = blockIdx.x;
  const int t0 = 2 * ((b & 7) * per_xcd + (b >> 3));
  if (t0 >= J.nblocks) return;               // the whole block
  const int tile = t0 + (threadIdx.x >> 7);
  twin64_tile<H, 2>(J, min(tile, J.nblocks - 1), tile < J.nblocks);
}

// Ensemble: every FP64 group of every domain whose nk the register solver
// takes, in one launch; tiles in XCD-aware order over the whole table.
template <int H>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
void k_vdiff_twin64_ens(const VdJob<double> *__restrict__ jobs, VdLaunch L, int per_xcd) {
  const int b = blockIdx.x;
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= L.first[L.njobs]) return;         // the whole block
  int lo = 0, hi = L.njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.first[mid] <= t) lo = mid; else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  twin64_tile<H>(jobs[lo], t - L.first[lo]);
}

__global__ void k_set_ints(int *__restrict__ p, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- compute_tmask_limit, ocean_tracer.F90:2788-2892 -----------------------
constexpr int NL = 32;
struct LimTable {
  const double *T[NL];
  double *lim[NL];
  double mn[NL], mx[NL];
  int n;
};

__global__ void k_tmask_limit(KDom d, LimTable t, int it, int is, int ts_same,
                              const int *__restrict__ kmt) {
  // 64 x 4 points of one level a work-group
  const int ii = blockIdx.x * 64 + threadIdx.x;   // gi = ii - 1
  const int gj = blockIdx.y * 4 + threadIdx.y - 1;
  const int k = blockIdx.z + 1;
  const int gi = ii - 1;
  if (gi > d.ni - 1 || gj > d.nj - 1) return;
  const int o = d2(d, gi, gj), P2 = d.P2, ok = o + P2 * (k - 1);
  const double m0 = (k <= kmt[o]) ? 1.0 : 0.0;
  const double me = (k <= kmt[o + 1]) ? 1.0 : 0.0;
  const double mn_ = (k <= kmt[o + d.NI]) ? 1.0 : 0.0;
  const double mb = (k < d.nk && k + 1 <= kmt[o]) ? 1.0 : 0.0;
  double flagT = 0.0, flagS = 0.0;
  for (int n = 0; n < t.n; ++n) {
    const double *T = t.T[n];
    const double lo = t.mn[n], hi = t.mx[n];
    double L = 0.0;
    double x = T[ok];
    if (m0 == 1.0 && (x < lo || x > hi)) L = m0;
    x = T[ok + 1];
    if (me == 1.0 && (x < lo || x > hi)) L = m0;
    x = T[ok + d.NI];
    if (mn_ == 1.0 && (x < lo || x > hi)) L = m0;
    if (k < d.nk) {
      x = T[ok + P2];
      if (mb == 1.0 && (x < lo || x > hi)) L = m0;
    }
    if (n == it) flagT = L;
    if (n == is) flagS = L;
    t.lim[n][ok] = L;
  }
  if (ts_same && it >= 0 && is >= 0 && (flagT == 1.0 || flagS == 1.0)) {
    t.lim[it][ok] = 1.0;
    t.lim[is][ok] = 1.0;
  }
}

// ---- explicit update, ocean_tracer.F90:2341-2350 ---------------------------
__global__ void k_explicit(KDom d, const double *__restrict__ Tm1, const double *__restrict__ th,
                           const double *__restrict__ rdz, const double *__restrict__ rdzr,
                           double dtime, double *__restrict__ Tp1) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  const int gj = blockIdx.y;
  const int k = blockIdx.z;
  if (gi >= d.ni) return;
  const int o = d2(d, gi, gj) + d.P2 * k;
  Tp1[o] = (rdz[o] * Tm1[o] + dtime * th[o]) * rdzr[o];
}

// ---- masked content with double-double accumulation -------------------------
struct DD {
  double hi, lo;
};
__device__ inline DD dd_add(DD a, double b) {
  const double s = a.hi + b;
  const double bb = s - a.hi;
  const double err = (a.hi - (s - bb)) + (b - bb);
  const double lo = a.lo + err;
  const double hi = s + lo;
  return {hi, lo - (hi - s)};
}
__device__ inline DD dd_add2(DD a, DD b) {
  DD r = dd_add(a, b.hi);
  return dd_add(r, b.lo);
}

constexpr int RB = 256;
template <typename TT>
__global__ __launch_bounds__(RB) void k_content(KDom d, const TT *__restrict__ T,
                                                const double *__restrict__ rho,
                                                const double *__restrict__ dat,
                                                const int *__restrict__ kmt,
                                                double *__restrict__ part) {
  __shared__ double sh[RB], sl[RB];
  DD acc{0.0, 0.0};
  const int ncol = d.ni * d.nj;
  for (int c = blockIdx.x * RB + threadIdx.x; c < ncol; c += gridDim.x * RB) {
    const int gi = c % d.ni, gj = c / d.ni;
    const int o = d2(d, gi, gj);
    const int km = kmt[o];
    const double a = dat[o];
    for (int k = 1; k <= km && k <= d.nk; ++k) {
      const int ok = o + d.P2 * (k - 1);
      acc = dd_add(acc, (double)T[ok] * (rho[ok] * a * 1.0));
    }
  }
  sh[threadIdx.x] = acc.hi;
  sl[threadIdx.x] = acc.lo;
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      DD x{sh[threadIdx.x], sl[threadIdx.x]}, y{sh[threadIdx.x + s], sl[threadIdx.x + s]};
      x = dd_add2(x, y);
      sh[threadIdx.x] = x.hi;
      sl[threadIdx.x] = x.lo;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0];
    part[2 * blockIdx.x + 1] = sl[0];
  }
}

__global__ __launch_bounds__(RB) void k_content_final(const double *__restrict__ part, int n,
                                                      double *__restrict__ out) {
  __shared__ double sh[RB], sl[RB];
  DD acc{0.0, 0.0};
  for (int b = threadIdx.x; b < n; b += RB) acc = dd_add2(acc, DD{part[2 * b], part[2 * b + 1]});
  sh[threadIdx.x] = acc.hi;
  sl[threadIdx.x] = acc.lo;
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      DD x{sh[threadIdx.x], sl[threadIdx.x]}, y{sh[threadIdx.x + s], sl[threadIdx.x + s]};
      x = dd_add2(x, y);
      sh[threadIdx.x] = x.hi;
      sl[threadIdx.x] = x.lo;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0] + sl[0];
}

// ---- error of a tracer against an FP64 reference (wet compute cells) -------
// Per block: max|a-b|, max|b|, sum (a-b)^2, sum b^2, cells, cells whose bits
// differ, max |a-b|/|b| where |b| >= floor.
constexpr int NERR = 7;
template <typename TT>
__global__ __launch_bounds__(RB) void k_tracer_error(KDom d, const TT *__restrict__ a,
                                                     const double *__restrict__ b,
                                                     const int *__restrict__ kmt, double floor,
                                                     double *__restrict__ part) {
  __shared__ double sh[NERR][RB];
  double v[NERR] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int ncol = d.ni * d.nj;
  for (int c = blockIdx.x * RB + threadIdx.x; c < ncol; c += gridDim.x * RB) {
    const int gi = c % d.ni, gj = c / d.ni;
    const int o = d2(d, gi, gj);
    const int km = kmt[o] < d.nk ? kmt[o] : d.nk;
    for (int k = 1; k <= km; ++k) {
      const int ok = o + d.P2 * (k - 1);
      const double x = (double)a[ok], y = b[ok];
      const double e = fabs(x - y), ay = fabs(y);
      v[0] = fmax(v[0], e);
      v[1] = fmax(v[1], ay);
      v[2] += e * e;
      v[3] += y * y;
      v[4] += 1.0;
      v[5] += __double_as_longlong(x) != __double_as_longlong(y) ? 1.0 : 0.0;
      if (ay >= floor && ay > 0.0) v[6] = fmax(v[6], e / ay);
    }
  }
  for (int q = 0; q < NERR; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      for (int q = 0; q < NERR; ++q) {
        const double x = sh[q][threadIdx.x], y = sh[q][threadIdx.x + s];
        sh[q][threadIdx.x] = (q == 0 || q == 1 || q == 6) ? fmax(x, y) : x + y;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < NERR) part[NERR * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// ---- grid mask consistency: tmask == (k<=kmt), umask == (k<=kmu) ------------
// One count per data-domain column; the host adds them up.
__global__ void k_check_masks(KDom d, const double *__restrict__ tmask,
                              const double *__restrict__ umask, const int *__restrict__ kmt,
                              const int *__restrict__ kmu, int *__restrict__ bad) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  int mine = 0;
  for (int k = 1; k <= d.nk; ++k) {
    const double t = tmask[o + d.P2 * (k - 1)], u = umask[o + d.P2 * (k - 1)];
    if (t != ((k <= kmt[o]) ? 1.0 : 0.0)) ++mine;
    if (u != ((k <= kmu[o]) ? 1.0 : 0.0)) ++mine;
  }
  bad[o] = mine;
}

__global__ void k_fill(double *__restrict__ p, double v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

int launch_copy(mom_gpu_ctx *ctx, double *dst, const double *src, size_t n) {
  if (n == 0) return MOM_GPU_OK;
  const size_t nb = (n + 255) / 256;
  const unsigned grid = (unsigned)(nb < 65536 ? nb : 65536);
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, ctx->stream, dst, src, n);
  MOM_LAUNCHED(ctx, "k_copy");
  return MOM_GPU_OK;
}

// ---- single-PE halo update (mpp_update_domains, shared/mpp/include/
// mpp_do_update.h:186-288): one thread per halo point of the data domain
// and level; a point the update leaves alone (closed axis) is not written.
__global__ void k_update_domains(KDom d, double *__restrict__ a, int nz) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;     // halo point index
  const int nhalo = 2 * d.NI + 2 * d.nj;                   // halo rows + halo columns
  if (h >= nhalo) return;
  int ii, jj;
  if (h < 2 * d.NI) {              // rows jj = 0 and NJ-1, every column
    ii = h % d.NI;
    jj = (h < d.NI) ? 0 : d.NJ - 1;
  } else {                         // columns ii = 0 and NI-1 of compute rows
    const int c = h - 2 * d.NI;
    jj = 1 + c % d.nj;
    ii = (c < d.nj) ? 0 : d.NI - 1;
  }
  int si, sj;
  if (!halo_source(d, ii, jj, si, sj) || (si == ii && sj == jj)) return;
  const size_t o = (size_t)ii + (size_t)d.NI * jj, s = (size_t)si + (size_t)d.NI * sj;
  const int k = blockIdx.y;
  if (k < nz) a[o + (size_t)d.P2 * k] = a[s + (size_t)d.P2 * k];
}

int launch_update_domains(mom_gpu_ctx *ctx, double *a, int nz) {
  const KDom &d = ctx->kd;
  const int nhalo = 2 * d.NI + 2 * d.nj;
  hipLaunchKernelGGL(k_update_domains, dim3((nhalo + 255) / 256, nz), dim3(256), 0, ctx->stream,
                     d, a, nz);
  MOM_LAUNCHED(ctx, "k_update_domains");
  return MOM_GPU_OK;
}

int launch_fill(mom_gpu_ctx *ctx, double *p, double v, size_t n) {
  if (n == 0) return MOM_GPU_OK;
  const size_t nb = (n + 255) / 256;
  const unsigned grid = (unsigned)(nb < 65536 ? nb : 65536);
  hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, ctx->stream, p, v, n);
  MOM_LAUNCHED(ctx, "k_fill");
  return MOM_GPU_OK;
}

int launch_invtri(mom_gpu_ctx *ctx, double *z, const double *topbc, const double *botbc,
                  const double *dcb, double tdt, const int *kmz, const double *mask,
                  const double *dh, const double *dhw, double aidif) {
  const KDom &d = ctx->kd;
  hipLaunchKernelGGL(k_invtri, dim3((d.ni + 63) / 64, d.nj), dim3(64), 0, ctx->stream, d, z,
                     topbc, botbc, dcb, tdt, kmz, mask, dh, dhw, aidif, ctx->tri_e);
  MOM_LAUNCHED(ctx, "k_invtri");
  return MOM_GPU_OK;
}

static int twin_slots(int nk);
static void launch_twin64(mom_gpu_ctx *ctx, VdJob<double> J);

// FP64 tracers take the register-resident solver (k_vdiff_twin64) whenever nk
// is in its range (2..80), whatever the knobs, so that FP64 results never
// depend on them; the Thomas kernels (k_vdiff1, k_vdiff) serve other nk.
// Jobs of up to NT tracers hold both mixing classes (class 1 first, n2 the
// first of class 2), so that T and S share one launch and the tile's rho_dzt
// and dzwt come from L2 for the second; a tracer with its own K33 goes alone.
static int twin64_groups(mom_gpu_ctx *ctx, int ntr, void *const *T, const int *dtype, int want,
                         const double *const *stf, const double *const *btf,
                         const double *diff_cbt, const double *const *K33, int index_salt,
                         double rho0, const double *dh, const double *dhw, double dt,
                         double aidif) {
  const KDom &d = ctx->kd;
  const size_t p3 = (size_t)d.P2 * d.nk;
  const int nbx = (d.ni + 63) / 64;
  auto job = [&]() {
    VdJob<double> J{};
    J.d = d; J.dcbase = diff_cbt; J.dcbase2 = diff_cbt + p3; J.rho0 = rho0; J.kmt = ctx->kmt;
    J.dh = dh; J.dhw = dhw; J.tdt = dt; J.aidif = aidif; J.ework = ctx->tri_e;
    J.nbx = nbx; J.nblocks = nbx * d.nj;
    return J;
  };
  std::vector<int> c1, c2;
  for (int n = 0; n < ntr; ++n) {
    if ((dtype ? dtype[n] : 8) != want) continue;
    const int nm = (n == index_salt) ? 2 : 1;
    const double *k33 = (K33 && K33[n]) ? K33[n] : nullptr;
    if (k33) {
      VdJob<double> J = job();
      J.g.z[0] = (double *)T[n]; J.g.stf[0] = stf[n]; J.g.btf[0] = btf[n]; J.g.n = 1;
      J.g.k33 = k33;
      J.n2 = (nm == 2) ? 0 : 1;
      launch_twin64(ctx, J);
      MOM_LAUNCHED(ctx, "k_vdiff_twin64");
      continue;
    }
    (nm == 1 ? c1 : c2).push_back(n);
  }
  const int n1 = (int)c1.size();
  std::vector<int> all(c1);
  all.insert(all.end(), c2.begin(), c2.end());
  for (size_t at = 0; at < all.size(); at += NT) {
    const int cnt = (int)std::min((size_t)NT, all.size() - at);
    VdJob<double> J = job();
    for (int q = 0; q < cnt; ++q) {
      const int n = all[at + q];
      J.g.z[q] = (double *)T[n]; J.g.stf[q] = stf[n]; J.g.btf[q] = btf[n];
    }
    J.g.n = cnt;
    J.n2 = std::max(0, std::min(cnt, n1 - (int)at));
    launch_twin64(ctx, J);
    MOM_LAUNCHED(ctx, "k_vdiff_twin64");
  }
  return MOM_GPU_OK;
}

template <typename TT>
static int vdiff_groups(mom_gpu_ctx *ctx, int ntr, void *const *T, const int *dtype, int want,
                        const double *const *stf, const double *const *btf,
                        const double *diff_cbt, const double *const *K33, int index_salt,
                        double rho0, const double *dh, const double *dhw, double dt,
                        double aidif) {
  const KDom &d = ctx->kd;
  if constexpr (sizeof(TT) == 8) {
    if (twin_slots(d.nk) > 0)
      return twin64_groups(ctx, ntr, T, dtype, want, stf, btf, diff_cbt, K33, index_salt, rho0,
                           dh, dhw, dt, aidif);
  }
  const size_t p3 = (size_t)d.P2 * d.nk;
  // tracers with their own K33 go alone; the others are grouped by nmix
  for (int nmix = 1; nmix <= 2; ++nmix) {
    TriGroup<TT> g{};
    g.n = 0;
    auto flush = [&](const double *k33) -> int {
      if (g.n == 0) return MOM_GPU_OK;
      g.k33 = k33;
      const int nbx = (d.ni + 63) / 64, per_xcd = (nbx * d.nj + 7) / 8;
      VdJob<TT> J{d, g, diff_cbt + p3 * (nmix - 1), rho0, ctx->kmt, dh, dhw, dt, aidif,
                  ctx->tri_e, nbx, nbx * d.nj};
      if constexpr (sizeof(TT) == 8) {
        if (g.n == 1 && d.nk > VD_TAIL_DEEP)
          hipLaunchKernelGGL(k_vdiff1<VD_TAIL_DEEP>, dim3(8 * per_xcd), dim3(64), 0,
                             ctx->stream, (const VdJob<double> &)J, per_xcd);
        else if (g.n == 1 && d.nk > VD_TAIL)
          hipLaunchKernelGGL(k_vdiff1<VD_TAIL>, dim3(8 * per_xcd), dim3(64), 0, ctx->stream,
                             (const VdJob<double> &)J, per_xcd);
        else
          hipLaunchKernelGGL(k_vdiff<TT>, dim3(8 * per_xcd), dim3(64), 0, ctx->stream, J,
                             per_xcd);
      } else {
        hipLaunchKernelGGL(k_vdiff<TT>, dim3(8 * per_xcd), dim3(64), 0, ctx->stream, J, per_xcd);
      }
      MOM_LAUNCHED(ctx, "k_vdiff");
      g.n = 0;
      return MOM_GPU_OK;
    };
    for (int n = 0; n < ntr; ++n) {
      if ((dtype ? dtype[n] : 8) != want) continue;
      const int nm = (n == index_salt) ? 2 : 1;
      if (nm != nmix) continue;
      const double *k33 = (K33 && K33[n]) ? K33[n] : nullptr;
      if (k33) {
        int rc = flush(nullptr);
        if (rc) return rc;
        g.z[0] = (TT *)T[n]; g.stf[0] = stf[n]; g.btf[0] = btf[n]; g.n = 1;
        rc = flush(k33);
        if (rc) return rc;
        continue;
      }
      g.z[g.n] = (TT *)T[n]; g.stf[g.n] = stf[n]; g.btf[g.n] = btf[n]; ++g.n;
      if (g.n == NT) {
        int rc = flush(nullptr);
        if (rc) return rc;
      }
    }
    int rc = flush(nullptr);
    if (rc) return rc;
  }
  return MOM_GPU_OK;
}

int launch_vdiff_implicit(mom_gpu_ctx *ctx, int ntr, double *const *T, const double *const *stf,
                          const double *const *btf, const double *diff_cbt,
                          const double *const *K33, int index_salt, double rho0,
                          const double *dh, const double *dhw, double dt, double aidif) {
  return vdiff_groups<double>(ctx, ntr, (void *const *)T, nullptr, 8, stf, btf, diff_cbt, K33,
                              index_salt, rho0, dh, dhw, dt, aidif);
}

// FP32 solver scratch of ctx: one plane per tracer of the largest group.
// Grows (outside graph capture only) when a call brings a larger group.
static int ensure_vd_scr(mom_gpu_ctx *ctx, mom_gpu_ctx *err, hipStream_t s, int planes,
                         const char *what) {
  if (ctx->vd_scr && ctx->vd_scr_planes >= planes) return MOM_GPU_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  MOM_HIP(err, hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone)
    return mom_fail(err, MOM_GPU_ERR_ARG,
                    "%s: a step with %d FP32 tracers per solver group allocates its scratch; "
                    "run one such step before capturing a graph", what, planes);
  if (ctx->vd_scr) {
    MOM_HIP(err, hipDeviceSynchronize());
    MOM_HIP(err, hipFree(ctx->vd_scr));
    ctx->vd_scr = nullptr;
    ctx->vd_scr_planes = 0;
  }
  const size_t p3 = (size_t)ctx->kd.P2 * ctx->kd.nk;
  MOM_HIP(err, hipMalloc((void **)&ctx->vd_scr, sizeof(float) * (size_t)planes * p3));
  ctx->vd_scr_planes = planes;
  return MOM_GPU_OK;
}

// FP32 tracers of the largest solver group of a call (by mixing class).
static int fp32_group_planes(int ntr, const int *dtype, int index_salt) {
  int cnt[2] = {0, 0};
  for (int n = 0; n < ntr; ++n)
    if (dtype[n] == 4) ++cnt[n == index_salt ? 1 : 0];
  return std::min(NTF, std::max(cnt[0], cnt[1]));
}

// One launch for the whole mix when it fits k_vdiff_mix: FP32 tracers only of
// class 1 (at most NTF), at most NT64 FP64 tracers of each class.
static int vdiff_mix_one(mom_gpu_ctx *ctx, int ntr, void *const *T, const int *dtype,
                         const double *const *stf, const double *const *btf,
                         const double *diff_cbt, int index_salt, double rho0, const double *dh,
                         const double *dhw, double dt, double aidif, bool &done) {
  done = false;
  VdMixJob J;
  std::memset((void *)&J, 0, sizeof J);
  for (int n = 0; n < ntr; ++n) {
    const int cls = (n == index_salt) ? 2 : 1;
    if (dtype[n] == 4) {
      if (cls != 1 || J.g32.n == NTF) return MOM_GPU_OK;
      J.g32.z[J.g32.n] = (float *)T[n]; J.g32.stf[J.g32.n] = stf[n];
      J.g32.btf[J.g32.n] = btf[n]; ++J.g32.n;
    } else if (cls == 1) {
      if (J.n1 == NT64) return MOM_GPU_OK;
      J.z1[J.n1] = (double *)T[n]; J.stf1[J.n1] = stf[n]; J.btf1[J.n1] = btf[n]; ++J.n1;
    } else {
      if (J.n2 == NT64) return MOM_GPU_OK;
      J.z2[J.n2] = (double *)T[n]; J.stf2[J.n2] = stf[n]; J.btf2[J.n2] = btf[n]; ++J.n2;
    }
  }
  if (J.g32.n == 0) return MOM_GPU_OK;
  const KDom &d = ctx->kd;
  const size_t p3 = (size_t)d.P2 * d.nk;
  int rc = ensure_vd_scr(ctx, ctx, ctx->stream, J.g32.n, "step_tracers");
  if (rc) return rc;
  if (J.n2 > 0 && !ctx->tri_e2) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MOM_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return mom_fail(ctx, MOM_GPU_ERR_ARG,
                      "step_tracers: the first mixed step allocates a second factorisation "
                      "plane; run one step before capturing a graph");
    MOM_HIP(ctx, hipMalloc((void **)&ctx->tri_e2, p3 * sizeof(double)));
  }
  J.d = d;
  J.dcb1 = diff_cbt;
  J.dcb2 = diff_cbt + p3;
  J.rho0 = rho0; J.kmt = ctx->kmt; J.dh = dh; J.dhw = dhw; J.tdt = dt; J.aidif = aidif;
  J.ew1 = ctx->tri_e; J.ew2 = ctx->tri_e2; J.scr = ctx->vd_scr;
  const int bx = ctx->tune.vdiff_block;
  const dim3 grid((d.ni + bx - 1) / bx, d.nj), block(bx);
  // register arrays sized to the group (4, 8, 10 or 12 FP32 tracers)
  const int nb = J.g32.n <= 4 ? 4 : J.g32.n <= 8 ? 8 : J.g32.n <= 10 ? 10 : 12;
#define MOM_VDMIX(PFV, NV) hipLaunchKernelGGL((k_vdiff_mix<PFV, NV>), grid, block, 0, ctx->stream, J)
  if (ctx->tune.vdiff_prefetch) {
    if (nb == 4) MOM_VDMIX(true, 4);
    else if (nb == 8) MOM_VDMIX(true, 8);
    else if (nb == 10) MOM_VDMIX(true, 10);
    else MOM_VDMIX(true, 12);
  } else {
    if (nb == 4) MOM_VDMIX(false, 4);
    else if (nb == 8) MOM_VDMIX(false, 8);
    else if (nb == 10) MOM_VDMIX(false, 10);
    else MOM_VDMIX(false, 12);
  }
#undef MOM_VDMIX
  MOM_LAUNCHED(ctx, "k_vdiff_mix");
  done = true;
  return MOM_GPU_OK;
}

// k_vdiff_twin register slots per wavefront for nk levels, H >= ceil(nk/2)
// (0: nk outside 2..80, where k_vdiff_inc serves).
static int twin_slots(int nk) {
  if (nk < 2 || nk > 80) return 0;
  const int h = (nk + 1) / 2;
  return h <= 8 ? 8 : h <= 16 ? 16 : h <= 24 ? 24 : h <= 32 ? 32 : 40;
}

static void launch_twin64(mom_gpu_ctx *ctx, VdJob<double> J) {
  const KDom &d = J.d;
  const int per_xcd = ((J.nblocks + 1) / 2 + 7) / 8;   // tile pairs per XCD
  const dim3 grid(8 * per_xcd), block(256);
  switch (twin_slots(d.nk)) {
    case 8: hipLaunchKernelGGL(k_vdiff_twin64<8>, grid, block, 0, ctx->stream, J, per_xcd); break;
    case 16: hipLaunchKernelGGL(k_vdiff_twin64<16>, grid, block, 0, ctx->stream, J, per_xcd); break;
    case 24: hipLaunchKernelGGL(k_vdiff_twin64<24>, grid, block, 0, ctx->stream, J, per_xcd); break;
    case 32: hipLaunchKernelGGL(k_vdiff_twin64<32>, grid, block, 0, ctx->stream, J, per_xcd); break;
    default: hipLaunchKernelGGL(k_vdiff_twin64<40>, grid, block, 0, ctx->stream, J, per_xcd); break;
  }
}

static void launch_twin(mom_gpu_ctx *ctx, VdJobF J) {
  const KDom &d = J.d;
  J.nbx = (d.ni + 63) / 64;
  J.nblocks = J.nbx * d.nj;
  const int per_xcd = ((J.nblocks + 1) / 2 + 7) / 8;   // tile pairs per XCD
  const dim3 grid(8 * per_xcd), block(256);
  switch (twin_slots(d.nk)) {
    case 8: hipLaunchKernelGGL(k_vdiff_twin<8>, grid, block, 0, ctx->stream, J, per_xcd); break;
    case 16: hipLaunchKernelGGL(k_vdiff_twin<16>, grid, block, 0, ctx->stream, J, per_xcd); break;
    case 24: hipLaunchKernelGGL(k_vdiff_twin<24>, grid, block, 0, ctx->stream, J, per_xcd); break;
    case 32: hipLaunchKernelGGL(k_vdiff_twin<32>, grid, block, 0, ctx->stream, J, per_xcd); break;
    default: hipLaunchKernelGGL(k_vdiff_twin<40>, grid, block, 0, ctx->stream, J, per_xcd); break;
  }
}

bool vdiff_fp32_scratch(const mom_gpu_ctx *c) {
  return !(c->tune.vdiff_solver == 1 && twin_slots(c->kd.nk) > 0);
}

int launch_vdiff_mixed(mom_gpu_ctx *ctx, int ntr, void *const *T, const int *dtype,
                       const double *const *stf, const double *const *btf,
                       const double *diff_cbt, int index_salt, double rho0, const double *dh,
                       const double *dhw, double dt, double aidif) {
  const bool twin = ctx->tune.vdiff_solver == 1 && twin_slots(ctx->kd.nk) > 0;
  if (ctx->tune.vdiff_merge && !twin && twin_slots(ctx->kd.nk) == 0) {
    bool done = false;
    const int rc0 = vdiff_mix_one(ctx, ntr, T, dtype, stf, btf, diff_cbt, index_salt, rho0, dh,
                                  dhw, dt, aidif, done);
    if (rc0 || done) return rc0;
  }
  int rc = vdiff_groups<double>(ctx, ntr, T, dtype, 8, stf, btf, diff_cbt, nullptr, index_salt,
                                rho0, dh, dhw, dt, aidif);
  if (rc) return rc;
  const KDom &d = ctx->kd;
  const size_t p3 = (size_t)d.P2 * d.nk;
  bool any32 = false;
  for (int n = 0; n < ntr; ++n) any32 = any32 || dtype[n] == 4;
  if (!any32) return MOM_GPU_OK;
  if (!twin) {
    rc = ensure_vd_scr(ctx, ctx, ctx->stream, fp32_group_planes(ntr, dtype, index_salt),
                       "step_tracers");
    if (rc) return rc;
  }
  for (int nmix = 1; nmix <= 2; ++nmix) {
    TriGroupF g{};
    auto flush = [&]() {
      if (g.n == 0) return;
      VdJobF J{d, g, diff_cbt + p3 * (nmix - 1), rho0, ctx->kmt, dh, dhw, dt, aidif,
               ctx->tri_e, ctx->vd_scr, 0, 0};
      const int bx = ctx->tune.vdiff_block;
      if (twin)
        launch_twin(ctx, J);
      else if (ctx->tune.vdiff_prefetch)
        hipLaunchKernelGGL(k_vdiff_inc<true>, dim3((d.ni + bx - 1) / bx, d.nj), dim3(bx), 0,
                           ctx->stream, J);
      else
        hipLaunchKernelGGL(k_vdiff_inc<false>, dim3((d.ni + bx - 1) / bx, d.nj), dim3(bx), 0,
                           ctx->stream, J);
      g.n = 0;
    };
    for (int n = 0; n < ntr; ++n) {
      if (dtype[n] != 4 || ((n == index_salt) ? 2 : 1) != nmix) continue;
      g.z[g.n] = (float *)T[n]; g.stf[g.n] = stf[n]; g.btf[g.n] = btf[n]; ++g.n;
      if (g.n == NTF) flush();
    }
    flush();
  }
  MOM_LAUNCHED(ctx, twin ? "k_vdiff_twin" : "k_vdiff_inc");
  return MOM_GPU_OK;
}

int launch_tmask_limit(mom_gpu_ctx *ctx, int ntr, const double *const *T, const double *mn,
                       const double *mx, int it, int is, int ts_same, double *const *lim) {
  const KDom &d = ctx->kd;
  if (ntr > NL) return mom_fail(ctx, MOM_GPU_ERR_ARG, "compute_tmask_limit: at most %d tracers", NL);
  LimTable t{};
  t.n = ntr;
  for (int n = 0; n < ntr; ++n) {
    t.T[n] = T[n]; t.lim[n] = lim[n]; t.mn[n] = mn[n]; t.mx[n] = mx[n];
  }
  // i in isc-1..iec, j in jsc-1..jec
  hipLaunchKernelGGL(k_tmask_limit, dim3((d.ni + 1 + 63) / 64, (d.nj + 1 + 3) / 4, d.nk),
                     dim3(64, 4), 0, ctx->stream, d, t, it, is, ts_same, ctx->kmt);
  MOM_LAUNCHED(ctx, "k_tmask_limit");
  return MOM_GPU_OK;
}

int launch_explicit(mom_gpu_ctx *ctx, const double *Tm1, const double *th, const double *rdz,
                    const double *rdzr, double dtime, double *Tp1) {
  const KDom &d = ctx->kd;
  hipLaunchKernelGGL(k_explicit, dim3((d.ni + 255) / 256, d.nj, d.nk), dim3(256), 0, ctx->stream,
                     d, Tm1, th, rdz, rdzr, dtime, Tp1);
  MOM_LAUNCHED(ctx, "k_explicit");
  return MOM_GPU_OK;
}

int launch_tracer_error(mom_gpu_ctx *ctx, const void *a, int dtype, const double *b,
                        double floor, double *out) {
  const KDom &d = ctx->kd;
  const int nb = ctx->red_n;
  if (dtype == 4)
    hipLaunchKernelGGL(k_tracer_error<float>, dim3(nb), dim3(RB), 0, ctx->stream, d,
                       (const float *)a, b, ctx->kmt, floor, ctx->red);
  else
    hipLaunchKernelGGL(k_tracer_error<double>, dim3(nb), dim3(RB), 0, ctx->stream, d,
                       (const double *)a, b, ctx->kmt, floor, ctx->red);
  MOM_LAUNCHED(ctx, "k_tracer_error");
  std::vector<double> part((size_t)NERR * nb);
  MOM_HIP(ctx, hipMemcpyAsync(part.data(), ctx->red, sizeof(double) * part.size(),
                              hipMemcpyDeviceToHost, ctx->stream));
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < NERR; ++q) out[q] = 0.0;
  for (int blk = 0; blk < nb; ++blk)
    for (int q = 0; q < NERR; ++q) {
      const double x = part[(size_t)NERR * blk + q];
      out[q] = (q == 0 || q == 1 || q == 6) ? (x > out[q] ? x : out[q]) : out[q] + x;
    }
  return MOM_GPU_OK;
}

int launch_content_part(mom_gpu_ctx *ctx, const void *T, int dtype, const double *rho,
                        double *part, double *out_dev) {
  const KDom &d = ctx->kd;
  const int nb = ctx->red_n;
  if (dtype == 4)
    hipLaunchKernelGGL(k_content<float>, dim3(nb), dim3(RB), 0, ctx->stream, d,
                       (const float *)T, rho, ctx->dat, ctx->kmt, part);
  else
    hipLaunchKernelGGL(k_content<double>, dim3(nb), dim3(RB), 0, ctx->stream, d,
                       (const double *)T, rho, ctx->dat, ctx->kmt, part);
  MOM_LAUNCHED(ctx, "k_content");
  hipLaunchKernelGGL(k_content_final, dim3(1), dim3(RB), 0, ctx->stream, part, nb, out_dev);
  MOM_LAUNCHED(ctx, "k_content_final");
  return MOM_GPU_OK;
}

int launch_content(mom_gpu_ctx *ctx, const void *T, int dtype, const double *rho, double *out) {
  const int nb = ctx->red_n;
  const int rc = launch_content_part(ctx, T, dtype, rho, ctx->red, ctx->red + 2 * nb);
  if (rc != MOM_GPU_OK) return rc;
  MOM_HIP(ctx, hipMemcpyAsync(out, ctx->red + 2 * nb, sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MOM_GPU_OK;
}

int launch_check_masks(mom_gpu_ctx *ctx, const double *tmask, const double *umask,
                       long long *bad) {
  const KDom &d = ctx->kd;
  const size_t n = (size_t)d.P2;
  int *dev = nullptr;
  MOM_HIP(ctx, hipMalloc(&dev, n * sizeof(int)));
  hipLaunchKernelGGL(k_check_masks, dim3((d.NI + 127) / 128, d.NJ), dim3(128), 0, ctx->stream, d,
                     tmask, umask, ctx->kmt, ctx->kmu, dev);
  hipError_t le = hipGetLastError();
  int *host = (int *)malloc(n * sizeof(int));
  hipError_t e1 = hipMemcpyAsync(host, dev, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dev);
  if (le != hipSuccess || e1 != hipSuccess || e2 != hipSuccess) {
    free(host);
    return mom_fail(ctx, MOM_GPU_ERR_HIP, "mask check failed: %s",
                    hipGetErrorString(le != hipSuccess ? le : (e1 != hipSuccess ? e1 : e2)));
  }
  long long s = 0;
  for (size_t q = 0; q < n; ++q) s += host[q];
  free(host);
  *bad = s;
  return MOM_GPU_OK;
}

int launch_set_ints(mom_gpu_ctx *ctx, int *p, int v, int n) {
  hipLaunchKernelGGL(k_set_ints, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, p, v, n);
  MOM_LAUNCHED(ctx, "k_set_ints");
  return MOM_GPU_OK;
}

// Implicit vertical diffusion of every domain of an ensemble: round r holds
// the r-th solver group of each domain, so no two jobs of a launch share a
// domain's scratch.
template <typename JOB, typename KER>
static int vd_rounds(mom_gpu_ensemble *e, std::vector<std::vector<JOB>> &groups, KER kernel,
                     const char *what) {
  size_t rounds = 0;
  for (auto &g : groups) rounds = std::max(rounds, g.size());
  for (size_t r = 0; r < rounds; ++r) {
    std::vector<unsigned char> buf;
    VdLaunch L{};
    int total = 0, nj = 0;
    for (auto &g : groups) {
      if (g.size() <= r) continue;
      const JOB &J = g[r];
      const size_t at = buf.size();
      buf.resize(at + sizeof(JOB));
      std::memcpy(buf.data() + at, &J, sizeof(JOB));
      L.first[nj++] = total;
      total += J.nblocks;
    }
    L.first[nj] = total;
    L.njobs = nj;
    if (total == 0) continue;
    const void *dev = ens_table(e, buf.data(), buf.size());
    if (!dev) return MOM_GPU_ERR_HIP;
    hipLaunchKernelGGL(kernel, dim3(total), dim3(64), 0, e->lead->stream, (const JOB *)dev, L);
    MOM_LAUNCHED(e->lead, what);
  }
  return MOM_GPU_OK;
}

int launch_vdiff_ensemble(mom_gpu_ensemble *e, const mom_gpu_tracer_step_t *steps) {
  const size_t nd = e->ctx.size();
  if (nd > (size_t)VD_MAXJ) return mom_fail(e->lead, MOM_GPU_ERR_ARG, "ensemble: too many domains");
  // the lead context's vdiff_solver chooses for every domain; the register
  // solver needs every domain's nk within its range
  bool twin = e->lead->tune.vdiff_solver == 1;
  int hmax = 0;
  for (size_t dd = 0; dd < nd; ++dd) {
    const int h = twin_slots(e->ctx[dd]->kd.nk);
    twin = twin && h > 0;
    hmax = std::max(hmax, h);
  }
  std::vector<std::vector<VdJob<double>>> g64(nd);
  std::vector<std::vector<VdJobF>> g32(nd);
  for (size_t dd = 0; dd < nd; ++dd) {
    const mom_gpu_tracer_step_t &s = steps[dd];
    if (!(s.stf && s.btf && s.aidif != 0.0) || s.ntr == 0) continue;
    mom_gpu_ctx *ctx = e->ctx[dd];
    const KDom &d = ctx->kd;
    const size_t p3 = (size_t)d.P2 * d.nk;
    const double *dh = s.rho_dzt_taup1 ? s.rho_dzt_taup1 : s.rho_dzt_tau;
    const int nbx = (d.ni + 63) / 64;
    bool any32 = false;
    for (int n = 0; n < s.ntr; ++n) any32 = any32 || s.dtype[n] == 4;
    if (any32 && !twin) {
      const int rc = ensure_vd_scr(ctx, e->lead, e->lead->stream,
                                   std::min(NTF_ENS, fp32_group_planes(s.ntr, s.dtype, s.index_salt)),
                                   "ensemble");
      if (rc) return rc;
    }
    for (int nmix = 1; nmix <= 2; ++nmix) {
      const double *dcb = s.diff_cbt + p3 * (nmix - 1);
      VdJob<double> j64;
      VdJobF j32;
      std::memset((void *)&j64, 0, sizeof j64);
      std::memset((void *)&j32, 0, sizeof j32);
      auto flush64 = [&]() {
        if (j64.g.n == 0) return;
        j64.d = d; j64.dcbase = dcb; j64.rho0 = s.rho0; j64.kmt = ctx->kmt; j64.dh = dh;
        j64.dhw = s.dzwt; j64.tdt = s.dtime; j64.aidif = s.aidif; j64.ework = ctx->tri_e;
        j64.nbx = nbx; j64.nblocks = nbx * d.nj;
        g64[dd].push_back(j64);
        std::memset((void *)&j64, 0, sizeof j64);
      };
      auto flush32 = [&]() {
        if (j32.g.n == 0) return;
        j32.d = d; j32.dcbase = dcb; j32.rho0 = s.rho0; j32.kmt = ctx->kmt; j32.dh = dh;
        j32.dhw = s.dzwt; j32.tdt = s.dtime; j32.aidif = s.aidif; j32.ework = ctx->tri_e;
        j32.scr = ctx->vd_scr; j32.nbx = nbx; j32.nblocks = nbx * d.nj;
        g32[dd].push_back(j32);
        std::memset((void *)&j32, 0, sizeof j32);
      };
      for (int n = 0; n < s.ntr; ++n) {
        if (((n == s.index_salt) ? 2 : 1) != nmix) continue;
        if (s.dtype[n] == 8) {
          TriGroup<double> &g = j64.g;
          g.z[g.n] = (double *)s.T_taup1[n]; g.stf[g.n] = s.stf[n]; g.btf[g.n] = s.btf[n];
          if (++g.n == NT) flush64();
        } else {
          TriGroupF &g = j32.g;
          g.z[g.n] = (float *)s.T_taup1[n]; g.stf[g.n] = s.stf[n]; g.btf[g.n] = s.btf[n];
          if (++g.n == (twin ? NTF : NTF_ENS)) flush32();
        }
      }
      flush64();
      flush32();
    }
  }
  // FP64: the register solver for every domain whose nk it takes (all their
  // groups in one launch, nothing shared), rounds of the Thomas kernel for
  // the others -- each domain as it would be solved alone
  {
    std::vector<VdJob<double>> all64;
    int h64 = 0;
    for (size_t dd = 0; dd < nd; ++dd) {
      const int h = twin_slots(e->ctx[dd]->kd.nk);
      if (h == 0) continue;
      h64 = std::max(h64, h);
      all64.insert(all64.end(), g64[dd].begin(), g64[dd].end());
      g64[dd].clear();
    }
    for (size_t at = 0; at < all64.size(); at += VD_MAXJ) {
      const size_t nj = std::min((size_t)VD_MAXJ, all64.size() - at);
      VdLaunch L{};
      int total = 0;
      for (size_t j = 0; j < nj; ++j) {
        L.first[j] = total;
        total += all64[at + j].nblocks;
      }
      L.first[nj] = total;
      L.njobs = (int)nj;
      const void *dev = ens_table(e, all64.data() + at, sizeof(VdJob<double>) * nj);
      if (!dev) return MOM_GPU_ERR_HIP;
      const int per_xcd = (total + 7) / 8;
      const dim3 grid(8 * per_xcd), block(128);
      hipStream_t st = e->lead->stream;
      const VdJob<double> *jt = (const VdJob<double> *)dev;
      switch (h64) {
        case 8: hipLaunchKernelGGL(k_vdiff_twin64_ens<8>, grid, block, 0, st, jt, L, per_xcd); break;
        case 16: hipLaunchKernelGGL(k_vdiff_twin64_ens<16>, grid, block, 0, st, jt, L, per_xcd); break;
        case 24: hipLaunchKernelGGL(k_vdiff_twin64_ens<24>, grid, block, 0, st, jt, L, per_xcd); break;
        case 32: hipLaunchKernelGGL(k_vdiff_twin64_ens<32>, grid, block, 0, st, jt, L, per_xcd); break;
        default: hipLaunchKernelGGL(k_vdiff_twin64_ens<40>, grid, block, 0, st, jt, L, per_xcd); break;
      }
      MOM_LAUNCHED(e->lead, "k_vdiff_twin64_ens");
    }
  }
  int rc = vd_rounds(e, g64, k_vdiff_ens, "k_vdiff_ens");
  if (rc) return rc;
  if (!twin) return vd_rounds(e, g32, k_vdiff_inc_ens, "k_vdiff_inc_ens");
  // register-resident solver: all jobs at once, VD_MAXJ a launch
  std::vector<VdJobF> all;
  for (auto &g : g32) all.insert(all.end(), g.begin(), g.end());
  for (size_t at = 0; at < all.size(); at += VD_MAXJ) {
    const size_t nj = std::min((size_t)VD_MAXJ, all.size() - at);
    VdLaunch L{};
    int total = 0;
    for (size_t j = 0; j < nj; ++j) {
      L.first[j] = total;
      total += all[at + j].nblocks;
    }
    L.first[nj] = total;
    L.njobs = (int)nj;
    const void *dev = ens_table(e, all.data() + at, sizeof(VdJobF) * nj);
    if (!dev) return MOM_GPU_ERR_HIP;
    const int per_xcd = (total + 7) / 8;
    const dim3 grid(8 * per_xcd), block(128);
    hipStream_t st = e->lead->stream;
    switch (hmax) {
      case 8: hipLaunchKernelGGL(k_vdiff_twin_ens<8>, grid, block, 0, st, (const VdJobF *)dev, L, per_xcd); break;
      case 16: hipLaunchKernelGGL(k_vdiff_twin_ens<16>, grid, block, 0, st, (const VdJobF *)dev, L, per_xcd); break;
      case 24: hipLaunchKernelGGL(k_vdiff_twin_ens<24>, grid, block, 0, st, (const VdJobF *)dev, L, per_xcd); break;
      case 32: hipLaunchKernelGGL(k_vdiff_twin_ens<32>, grid, block, 0, st, (const VdJobF *)dev, L, per_xcd); break;
      default: hipLaunchKernelGGL(k_vdiff_twin_ens<40>, grid, block, 0, st, (const VdJobF *)dev, L, per_xcd); break;
    }
    MOM_LAUNCHED(e->lead, "k_vdiff_twin_ens");
  }
  return MOM_GPU_OK;
}
