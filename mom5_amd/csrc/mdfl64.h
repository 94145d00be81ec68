// FP64 MDFL-Sweby face arithmetic shared by the per-tracer kernel (k_mdfl)
// and the batched tracer step (k_step<double>), so that T and S come out of
// the two paths bit for bit equal.
//
// The limited flux of ocean_tracer_advect.F90:3877-3885 (sweby_limiter = 1)
//   flux = 0.5*((M+|M|)*(T_up + psiP*Rj) + (M-|M|)*(T_dn - psiM*Rj))*mask,
//   psi  = max(0, min(1, d0 + d1*theta, (1-c)/(eps+c)*theta)),
//   theta = R_{j-1}/(eps + Rj) for psiP, R_{j+1}/(eps + Rj) for psiM,
// is formed without its two divisions per tracer and face: psi*Rj is the
// limiter multiplied through by Rj,
//   Rj >= 0:  max(0, min(Rj, d0*Rj + d1*Rn, kk*Rn)),
//   Rj <  0:  min(0, max(Rj, d0*Rj + d1*Rn, kk*Rn)),
// (Rn = R_{j-1} or R_{j+1}, kk = (1-c)/(eps+c)), equal in exact arithmetic
// wherever |Rj| >> 1e-30 and within 1e-30 elsewhere.  The face terms
// (c, d0, d1, kk and the upwind weight) are tracer independent.  Of the two
// terms only the upwind one has a nonzero weight: 0.5*(M+|M|) and
// 0.5*(M-|M|) are exactly M and 0 for M >= 0, 0 and M for M < 0, so the flux
// is M*mask*(T_up + psiP*Rj) or M*mask*(T_dn - psiM*Rj), the same value as
// the two-term sum (up to the sign of a zero), with one limited term formed
// instead of two.  The remaining reciprocals (of rho_dzt,
// of the face's thickness sum times its width, of eps + c) use rcp64: the
// hardware estimate refined by two Newton-Raphson steps, within one ulp of
// the correctly rounded quotient.  Every use is a plain FP64 multiply, add or
// fused multiply-add written out explicitly, so the result does not depend
// on compiler contraction.
#pragma once

#include <hip/hip_runtime.h>

namespace mdfl64 {

constexpr double ONESIXTH = 1.0 / 6.0;   // ocean_parameters.F90:92
constexpr double EPS30 = 1.0e-30;        // the limiter's guard

// 1/x for normal x (here 1e-30 <= x <= 1e12): v_rcp_f64's estimate, then
// y <- y + y*(1 - x*y) twice (Newton-Raphson for the reciprocal).
__device__ __forceinline__ double rcp64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// Tracer-independent terms of one face.
struct Face {
  double wm;                // M*mask, the weight of the upwind term
  unsigned long long nsg;   // the sign bit where M < 0 (the term enters negated)
  bool pos;                 // M >= 0: upwind is the Tpos side
  double d0, d1, kk;
};

// massflux M through the face, Courant number c >= 0, face mask mab in {0, 1}
__device__ __forceinline__ Face face(double massflux, double cfl, double mab) {
  Face f;
  f.d0 = (2.0 - cfl) * (1.0 - cfl) * ONESIXTH;
  f.d1 = (1.0 - cfl * cfl) * ONESIXTH;
  f.kk = (1.0 - cfl) * rcp64(EPS30 + cfl);
  f.pos = massflux >= 0.0;
  f.wm = massflux * mab;
  f.nsg = f.pos ? 0ull : (1ull << 63);
  return f;
}

// x with its sign bit flipped where sgn (0 or the sign bit) is set: an
// exact multiplication by -1
__device__ __forceinline__ double flip(double x, unsigned long long sgn) {
  return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, x) ^ sgn);
}

// psi*Rj for the neighbouring difference Rn; d0a = d0*Rj, sgn the sign bit
// of Rj.  With s = -1 where Rj's
// sign bit is set and +1 elsewhere (exact sign flips),
//   psi*Rj = s*max(0, min(|Rj|, s*lin, s*kr)),
// which is max(0, min(Rj, lin, kr)) for Rj >= 0 and min(0, max(Rj, lin, kr))
// for Rj < 0 without a branch.
__device__ __forceinline__ double psi_r(const Face &f, double d0a, double Rj,
                                        unsigned long long sgn, double Rn) {
  const double lin = fma(f.d1, Rn, d0a);
  const double kr = f.kk * Rn;
  const double m = __builtin_fmin(flip(lin, sgn), flip(kr, sgn));
  const double t = __builtin_fmax(0.0, __builtin_fmin(fabs(Rj), m));
  return flip(t, sgn);
}

// The limited flux: Tpos is upstream for M > 0, Tneg for M < 0; Rjm, Rj, Rjp
// the masked differences, as the reference forms them (a dry face has Rj = 0
// and a zero weight).  Rjm enters the limiter of the M > 0 term, Rjp that of
// the M < 0 term.
__device__ __forceinline__ double flux(const Face &f, double Rjm, double Rj, double Rjp,
                                       double Tpos, double Tneg) {
  const double d0a = f.d0 * Rj;
  const unsigned long long sgn = __builtin_bit_cast(unsigned long long, Rj) & (1ull << 63);
  const double Rn = f.pos ? Rjm : Rjp;
  const double Tu = f.pos ? Tpos : Tneg;
  return f.wm * (Tu + flip(psi_r(f, d0a, Rj, sgn, Rn), f.nsg));
}

// Courant numbers: |w*dtime/rho_dzt| of a bottom face (dtr = dtime/rho_dzt)
// and |u*dtime*2/((rho_a + rho_b)*dx)| of a horizontal face
__device__ __forceinline__ double cfl_z(double w, double dtr) { return fabs(w * dtr); }
__device__ __forceinline__ double cfl_h(double u, double dtime, double rsum, double dx) {
  return fabs(u * (2.0 * dtime) * rcp64(rsum * dx));
}

// dtime*tmask*datr/rho_dzt, the factor of the x- and y-sweep updates
__device__ __forceinline__ double sweep_factor(double dtr, double mi, double datr) {
  return dtr * mi * datr;
}

// the advective th_tendency contribution rho_dzt*(T_swept - T)/dtime*tmask
// (minus advect_tracer_mdfl_sweby of ocean_tracer_advect.F90:4043-4045), rdt = 1/dtime
__device__ __forceinline__ double tendency(double r, double tr, double Tc, double rdt,
                                           double mi) {
  return r * (tr - Tc) * rdt * mi;
}

}  // namespace mdfl64
