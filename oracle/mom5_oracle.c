/*
 * mom5_oracle.c -- CPU restatement of the MOM5 ocean hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in mom5_amd/
 * and the "port" CPU baseline.  The product never links or calls it.
 *
 * Each function names the reference routine (file:line, paths relative to
 * src/mom5/ of mom-ocean/MOM5) whose loop structure and left-to-right
 * evaluation order it follows.  Build flags: -O2 -ffp-contract=off (no
 * fused multiply-add), matching how the reference's FP64 arithmetic rounds.
 *
 * Halo semantics: one processing element owns the whole domain, so
 * mpp_update_domains reduces to a cyclic wrap on the cyclic axes; halo
 * points on a closed axis keep whatever value they hold (zero for the
 * module work arrays, which the reference zeroes before use).
 */
#include "mom5_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Loops marked PAR_FOR have independent iterations (no carried values, each
 * writes its own points), so the all-core build (-fopenmp, see Makefile)
 * gives results identical to the serial one; without -fopenmp the pragma is
 * ignored. */
#define PAR_FOR _Pragma("omp parallel for schedule(static)")

#define ONESIXTH (1.0 / 6.0)     /* ocean_core/ocean_parameters.F90:92 */
#define EPS30 1.0e-30            /* literal used by the MDFL limiter */

/* Fortran MIN/MAX on non-NaN data; ternaries fix the signed-zero choice so
 * the HIP kernels (same ternaries) round identically. */
static inline double fmn(double a, double b) { return (b < a) ? b : a; }
static inline double fmx(double a, double b) { return (b > a) ? b : a; }

/* ---- index helpers -------------------------------------------------- */
#define NI (d->ied - d->isd + 1)
#define NJ (d->jed - d->jsd + 1)
#define P2 ((size_t)NI * (size_t)NJ)
#define I2(i, j) ((size_t)((i) - d->isd) + (size_t)NI * (size_t)((j) - d->jsd))
#define I3(i, j, k) (I2(i, j) + P2 * (size_t)((k) - 1))
#define I3W(i, j, k) (I2(i, j) + P2 * (size_t)(k))
/* halo-2 work arrays (isc-2:iec+2, jsc-2:jec+2, nk) */
#define NI2 (d->iec - d->isc + 5)
#define NJ2 (d->jec - d->jsc + 5)
#define H2(i, j, k) ((size_t)((i) - (d->isc - 2)) + \
    (size_t)NI2 * ((size_t)((j) - (d->jsc - 2)) + (size_t)NJ2 * (size_t)((k) - 1)))
/* compute-domain arrays (isc:iec, jsc:jec, nk) */
#define NIC (d->iec - d->isc + 1)
#define NJC (d->jec - d->jsc + 1)
#define C3(i, j, k) ((size_t)((i) - d->isc) + \
    (size_t)NIC * ((size_t)((j) - d->jsc) + (size_t)NJC * (size_t)((k) - 1)))
#define C2(i, j) ((size_t)((i) - d->isc) + (size_t)NIC * (size_t)((j) - d->jsc))

/* Halo update of an array on (isc-h:iec+h, jsc-h:jec+h, nz), single PE.
 * mpp_do_update exchanges E/W strips over the compute rows, N/S strips
 * over the compute columns, and corners only where both neighbours exist
 * (shared/mpp/include/mpp_do_update.h:186-288). */
void orc_halo_update(const orc_dom *d, int h, double *a, int nz, int flags) {
  const int ni = d->iec - d->isc + 1, nj = d->jec - d->jsc + 1;
  const int nx = ni + 2 * h, ny = nj + 2 * h;
#define HA(ii, jj, kk) a[(size_t)((ii) - (d->isc - h)) + (size_t)nx * \
    ((size_t)((jj) - (d->jsc - h)) + (size_t)ny * (size_t)(kk))]
  PAR_FOR
  for (int k = 0; k < nz; ++k) {
    if ((flags & 1) && d->cyclic_x) {
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int s = 1; s <= h; ++s) {
          HA(d->isc - s, j, k) = HA(d->iec - s + 1, j, k);
          HA(d->iec + s, j, k) = HA(d->isc + s - 1, j, k);
        }
    }
    if ((flags & 2) && d->cyclic_y) {
      for (int i = d->isc; i <= d->iec; ++i)
        for (int s = 1; s <= h; ++s) {
          HA(i, d->jsc - s, k) = HA(i, d->jec - s + 1, k);
          HA(i, d->jec + s, k) = HA(i, d->jsc + s - 1, k);
        }
    }
    if (flags == 3 && d->cyclic_x && d->cyclic_y) {
      for (int s = 1; s <= h; ++s)
        for (int t = 1; t <= h; ++t) {
          HA(d->isc - s, d->jsc - t, k) = HA(d->iec - s + 1, d->jec - t + 1, k);
          HA(d->iec + s, d->jsc - t, k) = HA(d->isc + s - 1, d->jec - t + 1, k);
          HA(d->isc - s, d->jec + t, k) = HA(d->iec - s + 1, d->jsc + t - 1, k);
          HA(d->iec + s, d->jec + t, k) = HA(d->isc + s - 1, d->jsc + t - 1, k);
        }
    }
  }
#undef HA
}

/* tmask_mdfl: tmask on the compute domain, halo-updated on Dom_mdfl
 * (ocean_tracer_advect.F90:1644-1691, mdfl_init). */
static double *make_tmask_h2(const orc_dom *d, const orc_grid *g) {
  const int nk = d->nk;
  double *tm = (double *)calloc((size_t)NI2 * NJ2 * nk, sizeof(double));
  PAR_FOR
  for (int k = 1; k <= nk; ++k)
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) tm[H2(i, j, k)] = g->tmask[I3(i, j, k)];
  orc_halo_update(d, 2, tm, nk, 3);
  return tm;
}

/* MDFL limiter: psi = (d0+d1*theta)*(1-lim) + max(0,min(min(1,d0+d1*theta),
 * (1-cfl)/(eps+cfl)*theta))*lim   (ocean_tracer_advect.F90:3877-3885). */
static inline double psi_mdfl(double d0, double d1, double theta, double cfl,
                              double lim) {
  double p = d0 + d1 * theta;
  return p * (1.0 - lim) +
         fmx(0.0, fmn(fmn(1.0, d0 + d1 * theta), (1.0 - cfl) / (EPS30 + cfl) * theta)) * lim;
}
/* advect_tracer_sweby_all limiter (:4161-4164), the lim=1 form only. */
static inline double psi_all(double d0, double d1, double theta, double cfl) {
  return fmx(0.0, fmn(fmn(1.0, d0 + d1 * theta), (1.0 - cfl) / (EPS30 + cfl) * theta));
}

/* ------------------------------------------------------------------ */
/* ocean_tracer_advect.F90:3806-4066 advect_tracer_mdfl_sweby          */
/* ------------------------------------------------------------------ */
void orc_advect_tracer_mdfl_sweby(const orc_dom *d, const orc_grid *g,
    const double *T, const double *uhrho_et, const double *vhrho_nt,
    const double *wrho_bt, const double *rho_dzt, double dtime,
    double lim, double *advect, double *flux_x, double *flux_y,
    double *flux_z) {
  const int nk = d->nk;
  const size_t n3 = P2 * (size_t)nk;
  double *tm = make_tmask_h2(d, g);
  double *tr = (double *)calloc((size_t)NI2 * NJ2 * nk, sizeof(double));
  double *fx = (double *)calloc(n3, sizeof(double));
  double *fy = (double *)calloc(n3, sizeof(double));
  double *ftp = (double *)calloc((size_t)NIC * NJC, sizeof(double));
  double *wkm1 = (double *)calloc((size_t)NIC * NJC, sizeof(double));

  /* z-sweep: flux through the bottom face of each cell, :3843-3909 */
  for (int k = 1; k <= nk; ++k) {
    PAR_FOR
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) tr[H2(i, j, k)] = T[I3(i, j, k)];
    const int kp1 = (k + 1 < nk) ? k + 1 : nk;
    const int kp2 = (k + 2 < nk) ? k + 2 : nk;
    const int km1 = (k - 1 > 1) ? k - 1 : 1;
    PAR_FOR
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        const double Rjp = (T[I3(i, j, km1)] - T[I3(i, j, k)]) * tm[H2(i, j, km1)] * tm[H2(i, j, k)];
        const double Rj = (T[I3(i, j, k)] - T[I3(i, j, kp1)]) * tm[H2(i, j, k)] * tm[H2(i, j, kp1)];
        const double Rjm = (T[I3(i, j, kp1)] - T[I3(i, j, kp2)]) * tm[H2(i, j, kp1)] * tm[H2(i, j, kp2)];
        const double w = wrho_bt[I3W(i, j, k)];
        const double massflux = g->dat[I2(i, j)] * w;
        const double cfl = fabs(w * dtime / rho_dzt[I3(i, j, k)]);
        const double d0 = (2.0 - cfl) * (1.0 - cfl) * ONESIXTH;
        const double d1 = (1.0 - cfl * cfl) * ONESIXTH;
        const double thetaP = Rjm / (EPS30 + Rj);
        const double thetaM = Rjp / (EPS30 + Rj);
        const double psiP = psi_mdfl(d0, d1, thetaP, cfl, lim);
        const double psiM = psi_mdfl(d0, d1, thetaM, cfl, lim);
        const double fbt = 0.5 * ((massflux + fabs(massflux)) * (T[I3(i, j, kp1)] + psiP * Rj) +
                                  (massflux - fabs(massflux)) * (T[I3(i, j, k)] - psiM * Rj)) *
                           tm[H2(i, j, kp1)] * tm[H2(i, j, k)];
        tr[H2(i, j, k)] = tr[H2(i, j, k)] +
            dtime / rho_dzt[I3(i, j, k)] *
                (g->datr[I2(i, j)] * (fbt - ftp[C2(i, j)]) + T[I3(i, j, k)] * (wkm1[C2(i, j)] - w));
        if (flux_z) flux_z[I3(i, j, k)] = fbt;
        ftp[C2(i, j)] = fbt;
      }
    PAR_FOR
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) wkm1[C2(i, j)] = wrho_bt[I3W(i, j, k)];
  }
  orc_halo_update(d, 2, tr, nk, 1);

  /* x-sweep: flux through the eastern face, :3916-3969 */
  PAR_FOR
  for (int k = 1; k <= nk; ++k) {
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc - 1; i <= d->iec; ++i) {
        const double Rjp = (tr[H2(i + 2, j, k)] - tr[H2(i + 1, j, k)]) * tm[H2(i + 2, j, k)] * tm[H2(i + 1, j, k)];
        const double Rj = (tr[H2(i + 1, j, k)] - tr[H2(i, j, k)]) * tm[H2(i + 1, j, k)] * tm[H2(i, j, k)];
        const double Rjm = (tr[H2(i, j, k)] - tr[H2(i - 1, j, k)]) * tm[H2(i, j, k)] * tm[H2(i - 1, j, k)];
        const double u = uhrho_et[I3(i, j, k)];
        const double massflux = g->dyte[I2(i, j)] * u;
        const double cfl = fabs(u * dtime * 2.0 /
                                ((rho_dzt[I3(i, j, k)] + rho_dzt[I3(i + 1, j, k)]) * g->dxte[I2(i, j)]));
        const double d0 = (2.0 - cfl) * (1.0 - cfl) * ONESIXTH;
        const double d1 = (1.0 - cfl * cfl) * ONESIXTH;
        const double thetaP = Rjm / (EPS30 + Rj);
        const double thetaM = Rjp / (EPS30 + Rj);
        const double psiP = psi_mdfl(d0, d1, thetaP, cfl, lim);
        const double psiM = psi_mdfl(d0, d1, thetaM, cfl, lim);
        fx[I3(i, j, k)] = 0.5 * ((massflux + fabs(massflux)) * (tr[H2(i, j, k)] + psiP * Rj) +
                                 (massflux - fabs(massflux)) * (tr[H2(i + 1, j, k)] - psiM * Rj)) *
                          tm[H2(i, j, k)] * tm[H2(i + 1, j, k)];
      }
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        tr[H2(i, j, k)] = tr[H2(i, j, k)] +
            dtime * tm[H2(i, j, k)] * g->datr[I2(i, j)] / rho_dzt[I3(i, j, k)] *
                (fx[I3(i - 1, j, k)] - fx[I3(i, j, k)] +
                 T[I3(i, j, k)] * (g->dyte[I2(i, j)] * uhrho_et[I3(i, j, k)] -
                                   g->dyte[I2(i - 1, j)] * uhrho_et[I3(i - 1, j, k)]));
      }
  }
  orc_halo_update(d, 2, tr, nk, 2);

  /* y-sweep, then remove the divergence correction, :3974-4056 */
  memset(wkm1, 0, sizeof(double) * (size_t)NIC * NJC);
  for (int k = 1; k <= nk; ++k) {
    PAR_FOR
    for (int j = d->jsc - 1; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        const double Rjp = (tr[H2(i, j + 2, k)] - tr[H2(i, j + 1, k)]) * tm[H2(i, j + 2, k)] * tm[H2(i, j + 1, k)];
        const double Rj = (tr[H2(i, j + 1, k)] - tr[H2(i, j, k)]) * tm[H2(i, j + 1, k)] * tm[H2(i, j, k)];
        const double Rjm = (tr[H2(i, j, k)] - tr[H2(i, j - 1, k)]) * tm[H2(i, j, k)] * tm[H2(i, j - 1, k)];
        const double v = vhrho_nt[I3(i, j, k)];
        const double massflux = g->dxtn[I2(i, j)] * v;
        const double cfl = fabs(v * dtime * 2.0 /
                                ((rho_dzt[I3(i, j, k)] + rho_dzt[I3(i, j + 1, k)]) * g->dytn[I2(i, j)]));
        const double d0 = (2.0 - cfl) * (1.0 - cfl) * ONESIXTH;
        const double d1 = (1.0 - cfl * cfl) * ONESIXTH;
        const double thetaP = Rjm / (EPS30 + Rj);
        const double thetaM = Rjp / (EPS30 + Rj);
        const double psiP = psi_mdfl(d0, d1, thetaP, cfl, lim);
        const double psiM = psi_mdfl(d0, d1, thetaM, cfl, lim);
        fy[I3(i, j, k)] = 0.5 * ((massflux + fabs(massflux)) * (tr[H2(i, j, k)] + psiP * Rj) +
                                 (massflux - fabs(massflux)) * (tr[H2(i, j + 1, k)] - psiM * Rj)) *
                          tm[H2(i, j, k)] * tm[H2(i, j + 1, k)];
      }
    PAR_FOR
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i)
        tr[H2(i, j, k)] = tr[H2(i, j, k)] +
            dtime * tm[H2(i, j, k)] * g->datr[I2(i, j)] / rho_dzt[I3(i, j, k)] *
                (fy[I3(i, j - 1, k)] - fy[I3(i, j, k)]);
    PAR_FOR
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        const double Tc = T[I3(i, j, k)];
        const double rho = rho_dzt[I3(i, j, k)];
        tr[H2(i, j, k)] = tr[H2(i, j, k)] +
            dtime * Tc / rho *
                (wrho_bt[I3W(i, j, k)] - wkm1[C2(i, j)] +
                 g->datr[I2(i, j)] * (g->dyte[I2(i - 1, j)] * uhrho_et[I3(i - 1, j, k)] -
                                      g->dyte[I2(i, j)] * uhrho_et[I3(i, j, k)]));
        advect[C3(i, j, k)] = -(rho * (tr[H2(i, j, k)] - Tc) / dtime * tm[H2(i, j, k)]);
      }
    PAR_FOR
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) wkm1[C2(i, j)] = wrho_bt[I3W(i, j, k)];
  }
  if (flux_x) memcpy(flux_x, fx, n3 * sizeof(double));
  if (flux_y) memcpy(flux_y, fy, n3 * sizeof(double));
  free(tm); free(tr); free(fx); free(fy); free(ftp); free(wkm1);
}

/* ocean_tracer_advect.F90:1925-1996: wrk1 zeroed on the data domain,
 * wrk1(comp) = -advect, th_tendency(comp) += wrk1. */
void orc_horz_advect_tracer_mdfl(const orc_dom *d, const orc_grid *g,
    const double *T, const double *uhrho_et, const double *vhrho_nt,
    const double *wrho_bt, const double *rho_dzt, double dtime, double lim,
    double *th_tendency, double *wrk1) {
  const int nk = d->nk;
  double *adv = (double *)malloc((size_t)NIC * NJC * nk * sizeof(double));
  orc_advect_tracer_mdfl_sweby(d, g, T, uhrho_et, vhrho_nt, wrho_bt, rho_dzt, dtime, lim,
                               adv, NULL, NULL, NULL);
  if (wrk1) memset(wrk1, 0, P2 * (size_t)nk * sizeof(double));
  PAR_FOR
  for (int k = 1; k <= nk; ++k)
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        const double w1 = -adv[C3(i, j, k)];
        if (wrk1) wrk1[I3(i, j, k)] = w1;
        th_tendency[I3(i, j, k)] = th_tendency[I3(i, j, k)] + w1;
      }
  free(adv);
}

/* ------------------------------------------------------------------ */
/* ocean_tracer_advect.F90:4104-4511 advect_tracer_sweby_all            */
/* ------------------------------------------------------------------ */
void orc_advect_tracer_sweby_all(const orc_dom *d, const orc_grid *g, int ntr,
    const double *const *Tall, const double *uhrho_et, const double *vhrho_nt,
    const double *wrho_bt, const double *rho_dzt, double dtime,
    double *const *th_tendency, double *const *wrk1_out) {
  const int nk = d->nk;
  const size_t n3 = P2 * (size_t)nk;
  const size_t nh2 = (size_t)NI2 * NJ2 * nk;
  double *tm = make_tmask_h2(d, g);
  double *trall = (double *)calloc(nh2 * (size_t)ntr, sizeof(double));
  double *fx = (double *)calloc(n3, sizeof(double));
  double *fy = (double *)calloc(n3, sizeof(double));
  double *ftp = (double *)calloc((size_t)NIC * NJC, sizeof(double));
  double *wkm1 = (double *)calloc((size_t)NIC * NJC, sizeof(double));

  for (int n = 0; n < ntr; ++n) {
    const double *T = Tall[n];
    double *tr = trall + nh2 * (size_t)n;
    memset(ftp, 0, sizeof(double) * (size_t)NIC * NJC);
    memset(wkm1, 0, sizeof(double) * (size_t)NIC * NJC);
    if (wrk1_out) memset(wrk1_out[n], 0, n3 * sizeof(double));
    for (int k = 1; k <= nk; ++k) {
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int i = d->isc; i <= d->iec; ++i) tr[H2(i, j, k)] = T[I3(i, j, k)];
      const int kp1 = (k + 1 < nk) ? k + 1 : nk;
      const int kp2 = (k + 2 < nk) ? k + 2 : nk;
      const int km1 = (k - 1 > 1) ? k - 1 : 1;
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int i = d->isc; i <= d->iec; ++i) {
          const double Rjp = (T[I3(i, j, km1)] - T[I3(i, j, k)]) * tm[H2(i, j, km1)] * tm[H2(i, j, k)];
          const double Rj = (T[I3(i, j, k)] - T[I3(i, j, kp1)]) * tm[H2(i, j, k)] * tm[H2(i, j, kp1)];
          const double Rjm = (T[I3(i, j, kp1)] - T[I3(i, j, kp2)]) * tm[H2(i, j, kp1)] * tm[H2(i, j, kp2)];
          const double w = wrho_bt[I3W(i, j, k)];
          const double massflux = g->dat[I2(i, j)] * w;
          const double cfl = fabs(w * dtime / rho_dzt[I3(i, j, k)]);
          const double d0 = (2.0 - cfl) * (1.0 - cfl) * ONESIXTH;
          const double d1 = (1.0 - cfl * cfl) * ONESIXTH;
          const double thetaP = Rjm / (EPS30 + Rj);
          const double thetaM = Rjp / (EPS30 + Rj);
          const double psiP = psi_all(d0, d1, thetaP, cfl);
          const double psiM = psi_all(d0, d1, thetaM, cfl);
          const double fbt = 0.5 * ((massflux + fabs(massflux)) * (T[I3(i, j, kp1)] + psiP * Rj) +
                                    (massflux - fabs(massflux)) * (T[I3(i, j, k)] - psiM * Rj)) *
                             tm[H2(i, j, kp1)] * tm[H2(i, j, k)];
          const double wk = g->datr[I2(i, j)] * (fbt - ftp[C2(i, j)]) +
                            T[I3(i, j, k)] * (wkm1[C2(i, j)] - w);
          tr[H2(i, j, k)] = tr[H2(i, j, k)] + wk * dtime / rho_dzt[I3(i, j, k)];
          ftp[C2(i, j)] = fbt;
        }
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int i = d->isc; i <= d->iec; ++i) wkm1[C2(i, j)] = wrho_bt[I3W(i, j, k)];
    }
    orc_halo_update(d, 2, tr, nk, 1);
  }

  for (int n = 0; n < ntr; ++n) {
    const double *T = Tall[n];
    double *tr = trall + nh2 * (size_t)n;
    for (int k = 1; k <= nk; ++k) {
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int i = d->isc - 1; i <= d->iec; ++i) {
          const double Rjp = (tr[H2(i + 2, j, k)] - tr[H2(i + 1, j, k)]) * tm[H2(i + 2, j, k)] * tm[H2(i + 1, j, k)];
          const double Rj = (tr[H2(i + 1, j, k)] - tr[H2(i, j, k)]) * tm[H2(i + 1, j, k)] * tm[H2(i, j, k)];
          const double Rjm = (tr[H2(i, j, k)] - tr[H2(i - 1, j, k)]) * tm[H2(i, j, k)] * tm[H2(i - 1, j, k)];
          const double u = uhrho_et[I3(i, j, k)];
          const double massflux = g->dyte[I2(i, j)] * u;
          const double cfl = fabs(u * dtime * 2.0 /
                                  ((rho_dzt[I3(i, j, k)] + rho_dzt[I3(i + 1, j, k)]) * g->dxte[I2(i, j)]));
          const double d0 = (2.0 - cfl) * (1.0 - cfl) * ONESIXTH;
          const double d1 = (1.0 - cfl * cfl) * ONESIXTH;
          const double thetaP = Rjm / (EPS30 + Rj);
          const double thetaM = Rjp / (EPS30 + Rj);
          const double psiP = psi_all(d0, d1, thetaP, cfl);
          const double psiM = psi_all(d0, d1, thetaM, cfl);
          fx[I3(i, j, k)] = 0.5 * ((massflux + fabs(massflux)) * (tr[H2(i, j, k)] + psiP * Rj) +
                                   (massflux - fabs(massflux)) * (tr[H2(i + 1, j, k)] - psiM * Rj)) *
                            tm[H2(i, j, k)] * tm[H2(i + 1, j, k)];
        }
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int i = d->isc; i <= d->iec; ++i) {
          const double wk = tm[H2(i, j, k)] * g->datr[I2(i, j)] *
              (fx[I3(i - 1, j, k)] - fx[I3(i, j, k)] +
               T[I3(i, j, k)] * (g->dyte[I2(i, j)] * uhrho_et[I3(i, j, k)] -
                                 g->dyte[I2(i - 1, j)] * uhrho_et[I3(i - 1, j, k)]));
          tr[H2(i, j, k)] = tr[H2(i, j, k)] + wk * dtime / rho_dzt[I3(i, j, k)];
        }
    }
    orc_halo_update(d, 2, tr, nk, 2);
  }

  for (int n = 0; n < ntr; ++n) {
    const double *T = Tall[n];
    double *tr = trall + nh2 * (size_t)n;
    memset(wkm1, 0, sizeof(double) * (size_t)NIC * NJC);
    for (int k = 1; k <= nk; ++k) {
      for (int j = d->jsc - 1; j <= d->jec; ++j)
        for (int i = d->isc; i <= d->iec; ++i) {
          const double Rjp = (tr[H2(i, j + 2, k)] - tr[H2(i, j + 1, k)]) * tm[H2(i, j + 2, k)] * tm[H2(i, j + 1, k)];
          const double Rj = (tr[H2(i, j + 1, k)] - tr[H2(i, j, k)]) * tm[H2(i, j + 1, k)] * tm[H2(i, j, k)];
          const double Rjm = (tr[H2(i, j, k)] - tr[H2(i, j - 1, k)]) * tm[H2(i, j, k)] * tm[H2(i, j - 1, k)];
          const double v = vhrho_nt[I3(i, j, k)];
          const double massflux = g->dxtn[I2(i, j)] * v;
          const double cfl = fabs(v * dtime * 2.0 /
                                  ((rho_dzt[I3(i, j, k)] + rho_dzt[I3(i, j + 1, k)]) * g->dytn[I2(i, j)]));
          const double d0 = (2.0 - cfl) * (1.0 - cfl) * ONESIXTH;
          const double d1 = (1.0 - cfl * cfl) * ONESIXTH;
          const double thetaP = Rjm / (EPS30 + Rj);
          const double thetaM = Rjp / (EPS30 + Rj);
          const double psiP = psi_all(d0, d1, thetaP, cfl);
          const double psiM = psi_all(d0, d1, thetaM, cfl);
          fy[I3(i, j, k)] = 0.5 * ((massflux + fabs(massflux)) * (tr[H2(i, j, k)] + psiP * Rj) +
                                   (massflux - fabs(massflux)) * (tr[H2(i, j + 1, k)] - psiM * Rj)) *
                            tm[H2(i, j, k)] * tm[H2(i, j + 1, k)];
        }
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int i = d->isc; i <= d->iec; ++i) {
          const double Tc = T[I3(i, j, k)];
          const double rho = rho_dzt[I3(i, j, k)];
          const double wk = tm[H2(i, j, k)] * g->datr[I2(i, j)] * (fy[I3(i, j - 1, k)] - fy[I3(i, j, k)]) +
              Tc * (wrho_bt[I3W(i, j, k)] - wkm1[C2(i, j)] +
                    g->datr[I2(i, j)] * (g->dyte[I2(i - 1, j)] * uhrho_et[I3(i - 1, j, k)] -
                                         g->dyte[I2(i, j)] * uhrho_et[I3(i, j, k)]));
          tr[H2(i, j, k)] = tr[H2(i, j, k)] + wk * dtime / rho;
          const double w1 = rho * (tr[H2(i, j, k)] - Tc) / dtime * tm[H2(i, j, k)];
          if (wrk1_out) wrk1_out[n][I3(i, j, k)] = w1;
          th_tendency[n][I3(i, j, k)] = th_tendency[n][I3(i, j, k)] + w1;
        }
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int i = d->isc; i <= d->iec; ++i) wkm1[C2(i, j)] = wrho_bt[I3W(i, j, k)];
    }
  }
  free(tm); free(trall); free(fx); free(fy); free(ftp); free(wkm1);
}

/* ------------------------------------------------------------------ */
/* ocean_tracer_advect.F90:1442-1586 quicker_init                       */
/* ------------------------------------------------------------------ */
void orc_quicker_init(const orc_dom *d, const orc_grid *g, double *quick_x,
    double *quick_y, double *curv_xp, double *curv_xn, double *curv_yp,
    double *curv_yn, double *quick_z, double *curv_zp, double *curv_zn) {
  const int nk = d->nk;
  const size_t nq = (size_t)NI2 * NJ2;
  double *dxq = (double *)calloc(nq, sizeof(double));
  double *dyq = (double *)calloc(nq, sizeof(double));
#define Q2(i, j) H2(i, j, 1)
  for (int i = d->isc; i <= d->iec; ++i)
    for (int j = d->jsc; j <= d->jec; ++j) {
      dxq[Q2(i, j)] = g->dxt[I2(i, j)];
      dyq[Q2(i, j)] = g->dyt[I2(i, j)];
    }
  /* boundary fill (:1494-1514) */
  for (int i = d->isc - 2; i <= d->isc - 1; ++i) {
    for (int j = d->jsc; j <= d->jec; ++j) dxq[Q2(i, j)] = dxq[Q2(d->isc, j)];
    for (int j = d->jsc - 2; j <= d->jec + 2; ++j) dyq[Q2(i, j)] = dyq[Q2(d->isc, j)];
  }
  for (int i = d->iec + 1; i <= d->iec + 2; ++i) {
    for (int j = d->jsc; j <= d->jec; ++j) dxq[Q2(i, j)] = dxq[Q2(d->iec, j)];
    for (int j = d->jsc - 2; j <= d->jec + 2; ++j) dyq[Q2(i, j)] = dyq[Q2(d->iec, j)];
  }
  for (int j = d->jsc - 2; j <= d->jsc - 1; ++j)
    for (int i = d->isc - 2; i <= d->iec + 2; ++i) {
      dxq[Q2(i, j)] = dxq[Q2(i, d->jsc)];
      dyq[Q2(i, j)] = dyq[Q2(i, d->jsc)];
    }
  for (int j = d->jec + 1; j <= d->jec + 2; ++j)
    for (int i = d->isc - 2; i <= d->iec + 2; ++i) {
      dxq[Q2(i, j)] = dxq[Q2(i, d->jec)];
      dyq[Q2(i, j)] = dyq[Q2(i, d->jec)];
    }
  orc_halo_update(d, 2, dxq, 1, 3);
  orc_halo_update(d, 2, dyq, 1, 3);

  memset(quick_x, 0, 2 * P2 * sizeof(double));
  memset(quick_y, 0, 2 * P2 * sizeof(double));
  memset(curv_xp, 0, 3 * P2 * sizeof(double));
  memset(curv_xn, 0, 3 * P2 * sizeof(double));
  memset(curv_yp, 0, 3 * P2 * sizeof(double));
  memset(curv_yn, 0, 3 * P2 * sizeof(double));
  for (int i = d->isc - 1; i <= d->iec; ++i)
    for (int j = d->jsc - 1; j <= d->jec; ++j) {
      const double xm = dxq[Q2(i - 1, j)], x0 = dxq[Q2(i, j)], x1 = dxq[Q2(i + 1, j)], x2 = dxq[Q2(i + 2, j)];
      const double ym = dyq[Q2(i, j - 1)], y0 = dyq[Q2(i, j)], y1 = dyq[Q2(i, j + 1)], y2 = dyq[Q2(i, j + 2)];
      const size_t o = I2(i, j);
      quick_x[o] = x1 / (x1 + x0);
      quick_x[o + P2] = x0 / (x1 + x0);
      quick_y[o] = y1 / (y1 + y0);
      quick_y[o + P2] = y0 / (y1 + y0);
      curv_xp[o] = x0 * x1 / ((xm + 2.0 * x0 + x1) * (x0 + x1));
      curv_xp[o + P2] = -x0 * x1 / ((x0 + x1) * (xm + x0));
      curv_xp[o + 2 * P2] = x0 * x1 / ((xm + 2.0 * x0 + x1) * (xm + x0));
      curv_xn[o] = x0 * x1 / ((x0 + 2.0 * x1 + x2) * (x1 + x2));
      curv_xn[o + P2] = -x0 * x1 / ((x1 + x2) * (x0 + x1));
      curv_xn[o + 2 * P2] = x0 * x1 / ((x0 + 2.0 * x1 + x2) * (x0 + x1));
      curv_yp[o] = y0 * y1 / ((ym + 2.0 * y0 + y1) * (y0 + y1));
      curv_yp[o + P2] = -y0 * y1 / ((y0 + y1) * (ym + y0));
      curv_yp[o + 2 * P2] = y0 * y1 / ((ym + 2.0 * y0 + y1) * (ym + y0));
      curv_yn[o] = y0 * y1 / ((y0 + 2.0 * y1 + y2) * (y1 + y2));
      curv_yn[o + P2] = -y0 * y1 / ((y1 + y2) * (y0 + y1));
      curv_yn[o + 2 * P2] = y0 * y1 / ((y0 + 2.0 * y1 + y2) * (y0 + y1));
    }
#undef Q2
  for (int k = 1; k <= nk; ++k) {
    const int kp2 = (k + 2 < nk) ? k + 2 : nk;
    const int kp1 = (k + 1 < nk) ? k + 1 : nk;
    const int km1 = (k - 1 > 1) ? k - 1 : 1;
    const double zm = g->dzt[km1 - 1], z0 = g->dzt[k - 1], z1 = g->dzt[kp1 - 1], z2 = g->dzt[kp2 - 1];
    quick_z[k - 1] = z1 / (z1 + z0);
    quick_z[k - 1 + nk] = z0 / (z1 + z0);
    curv_zp[k - 1] = z0 * z1 / ((zm + 2.0 * z0 + z1) * (z0 + z1));
    curv_zp[k - 1 + nk] = -z0 * z1 / ((z0 + z1) * (zm + z0));
    curv_zp[k - 1 + 2 * nk] = z0 * z1 / ((zm + 2.0 * z0 + z1) * (zm + z0));
    curv_zn[k - 1] = z0 * z1 / ((z0 + 2.0 * z1 + z2) * (z1 + z2));
    curv_zn[k - 1 + nk] = -z0 * z1 / ((z1 + z2) * (z0 + z1));
    curv_zn[k - 1 + 2 * nk] = z0 * z1 / ((z0 + 2.0 * z1 + z2) * (z0 + z1));
  }
  free(dxq); free(dyq);
}

/* ------------------------------------------------------------------ */
/* ocean_tracer_advect.F90:2538-2653 horz_advect_tracer_quicker         */
/* ------------------------------------------------------------------ */
void orc_horz_advect_tracer_quicker(const orc_dom *d, const orc_grid *g,
    const double *Tm1, const double *Ttau, const double *uhrho_et,
    const double *vhrho_nt, const double *tmask_limit, int limit_with_upwind,
    double *result, double *flux_x_out, double *flux_y_out) {
  const int nk = d->nk;
  const size_t n3 = P2 * (size_t)nk;
  double *qx = (double *)malloc(2 * P2 * sizeof(double));
  double *qy = (double *)malloc(2 * P2 * sizeof(double));
  double *cxp = (double *)malloc(3 * P2 * sizeof(double));
  double *cxn = (double *)malloc(3 * P2 * sizeof(double));
  double *cyp = (double *)malloc(3 * P2 * sizeof(double));
  double *cyn = (double *)malloc(3 * P2 * sizeof(double));
  double *qz = (double *)malloc(2 * (size_t)nk * sizeof(double));
  double *czp = (double *)malloc(3 * (size_t)nk * sizeof(double));
  double *czn = (double *)malloc(3 * (size_t)nk * sizeof(double));
  orc_quicker_init(d, g, qx, qy, cxp, cxn, cyp, cyn, qz, czp, czn);
  double *tq = (double *)calloc((size_t)NI2 * NJ2 * nk, sizeof(double));
  double *tmq = make_tmask_h2(d, g);   /* tmask_quick is built like tmask_mdfl */
  double *fx = (double *)calloc(n3, sizeof(double));
  double *fy = (double *)calloc(n3, sizeof(double));
  PAR_FOR
  for (int k = 1; k <= nk; ++k)
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) tq[H2(i, j, k)] = Tm1[I3(i, j, k)];
  orc_halo_update(d, 2, tq, nk, 3);

  PAR_FOR
  for (int k = 1; k <= nk; ++k) {
    for (int j = d->jsc - 1; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        const size_t o = I2(i, j);
        const double vel = g->dxtn[o] * vhrho_nt[I3(i, j, k)];
        const double upos = 0.5 * (vel + fabs(vel));
        const double uneg = 0.5 * (vel - fabs(vel));
        const double rnormsk = tmq[H2(i, j, k)] * (1.0 - tmq[H2(i, j - 1, k)]);
        const double soutmsk = tmq[H2(i, j + 1, k)] * (1.0 - tmq[H2(i, j + 2, k)]);
        fy[I3(i, j, k)] =
            vel * (qy[o] * Ttau[I3(i, j, k)] + qy[o + P2] * Ttau[I3(i, j + 1, k)]) -
            upos * (cyp[o] * tq[H2(i, j + 1, k)] + cyp[o + P2] * tq[H2(i, j, k)] +
                    cyp[o + 2 * P2] * (tq[H2(i, j - 1, k)] * (1.0 - rnormsk) + tq[H2(i, j, k)] * rnormsk)) -
            uneg * (cyn[o] * (tq[H2(i, j + 2, k)] * (1.0 - soutmsk) + tq[H2(i, j + 1, k)] * soutmsk) +
                    cyn[o + P2] * tq[H2(i, j + 1, k)] + cyn[o + 2 * P2] * tq[H2(i, j, k)]);
      }
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc - 1; i <= d->iec; ++i) {
        const size_t o = I2(i, j);
        const double vel = g->dyte[o] * uhrho_et[I3(i, j, k)];
        const double upos = 0.5 * (vel + fabs(vel));
        const double uneg = 0.5 * (vel - fabs(vel));
        const double eastmsk = tmq[H2(i, j, k)] * (1.0 - tmq[H2(i - 1, j, k)]);
        const double westmsk = tmq[H2(i + 1, j, k)] * (1.0 - tmq[H2(i + 2, j, k)]);
        fx[I3(i, j, k)] =
            vel * (qx[o] * Ttau[I3(i, j, k)] + qx[o + P2] * Ttau[I3(i + 1, j, k)]) -
            upos * (cxp[o] * tq[H2(i + 1, j, k)] + cxp[o + P2] * tq[H2(i, j, k)] +
                    cxp[o + 2 * P2] * (tq[H2(i - 1, j, k)] * (1.0 - eastmsk) + tq[H2(i, j, k)] * eastmsk)) -
            uneg * (cxn[o] * (tq[H2(i + 2, j, k)] * (1.0 - westmsk) + tq[H2(i + 1, j, k)] * westmsk) +
                    cxn[o + P2] * tq[H2(i + 1, j, k)] + cxn[o + 2 * P2] * tq[H2(i, j, k)]);
      }
    if (limit_with_upwind) {
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int i = d->isc - 1; i <= d->iec; ++i)
          if (tmask_limit[I3(i, j, k)] == 1.0) {
            const double vel = uhrho_et[I3(i, j, k)];
            const double upos = 0.5 * (vel + fabs(vel));
            const double uneg = 0.5 * (vel - fabs(vel));
            fx[I3(i, j, k)] = g->dyte[I2(i, j)] * (upos * Tm1[I3(i, j, k)] + uneg * Tm1[I3(i + 1, j, k)]) *
                              g->tmask[I3(i, j, k)] * g->tmask[I3(i + 1, j, k)];
          }
      for (int j = d->jsc - 1; j <= d->jec; ++j)
        for (int i = d->isc; i <= d->iec; ++i)
          if (tmask_limit[I3(i, j, k)] == 1.0) {
            const double vel = vhrho_nt[I3(i, j, k)];
            const double upos = 0.5 * (vel + fabs(vel));
            const double uneg = 0.5 * (vel - fabs(vel));
            fy[I3(i, j, k)] = g->dxtn[I2(i, j)] * (upos * Tm1[I3(i, j, k)] + uneg * Tm1[I3(i, j + 1, k)]) *
                              g->tmask[I3(i, j, k)] * g->tmask[I3(i, j + 1, k)];
          }
    }
  }
  PAR_FOR
  for (int k = 1; k <= nk; ++k)
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i)
        result[C3(i, j, k)] = g->tmask[I3(i, j, k)] *
            (fx[I3(i, j, k)] - fx[I3(i - 1, j, k)] + fy[I3(i, j, k)] - fy[I3(i, j - 1, k)]) *
            g->datr[I2(i, j)];
  if (flux_x_out) memcpy(flux_x_out, fx, n3 * sizeof(double));
  if (flux_y_out) memcpy(flux_y_out, fy, n3 * sizeof(double));
  free(qx); free(qy); free(cxp); free(cxn); free(cyp); free(cyn);
  free(qz); free(czp); free(czn); free(tq); free(tmq); free(fx); free(fy);
}

/* ------------------------------------------------------------------ */
/* ocean_tracer_advect.F90:2981-3031 vert_advect_tracer_quicker         */
/* ------------------------------------------------------------------ */
void orc_vert_advect_tracer_quicker(const orc_dom *d, const orc_grid *g,
    const double *Tm1, const double *Ttau, const double *wrho_bt,
    const double *tmask_limit, double *result, double *flux_z) {
  const int nk = d->nk;
  double *qx = (double *)malloc(2 * P2 * sizeof(double));
  double *qy = (double *)malloc(2 * P2 * sizeof(double));
  double *cxp = (double *)malloc(3 * P2 * sizeof(double));
  double *cxn = (double *)malloc(3 * P2 * sizeof(double));
  double *cyp = (double *)malloc(3 * P2 * sizeof(double));
  double *cyn = (double *)malloc(3 * P2 * sizeof(double));
  double *qz = (double *)malloc(2 * (size_t)nk * sizeof(double));
  double *czp = (double *)malloc(3 * (size_t)nk * sizeof(double));
  double *czn = (double *)malloc(3 * (size_t)nk * sizeof(double));
  orc_quicker_init(d, g, qx, qy, cxp, cxn, cyp, cyn, qz, czp, czn);
  double *ft1 = (double *)calloc((size_t)NIC * NJC, sizeof(double));
  for (int k = 1; k <= nk; ++k) {
    const int km1 = (k - 1 > 1) ? k - 1 : 1;
    const int kp1 = (k + 1 < nk) ? k + 1 : nk;
    const int p2 = (k + 2 < nk) ? k + 2 : nk;
    PAR_FOR
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        const double vel = wrho_bt[I3W(i, j, k)];
        const double upos = 0.5 * (vel + fabs(vel));
        const double uneg = 0.5 * (vel - fabs(vel));
        const double *tmk = g->tmask;
        double ft2;
        if (tmask_limit[I3(i, j, k)] == 1.0) {
          ft2 = (uneg * Tm1[I3(i, j, k)] + upos * Tm1[I3(i, j, kp1)]) * tmk[I3(i, j, k)] * tmk[I3(i, j, kp1)];
        } else {
          const double tp2 = tmk[I3(i, j, p2)];
          const int kp2 = (int)lround(tp2 * (double)p2 + (1.0 - tp2) * (double)kp1);
          const double upmsk = tmk[I3(i, j, k)] * (1.0 - tmk[I3(i, j, km1)]);
          ft2 = vel * (qz[k - 1] * Ttau[I3(i, j, k)] + qz[k - 1 + nk] * Ttau[I3(i, j, kp1)]) -
                uneg * (czp[k - 1] * Tm1[I3(i, j, kp1)] + czp[k - 1 + nk] * Tm1[I3(i, j, k)] +
                        czp[k - 1 + 2 * nk] * Tm1[I3(i, j, km1)]) -
                upos * (czn[k - 1] * Tm1[I3(i, j, kp2)] + czn[k - 1 + nk] * Tm1[I3(i, j, kp1)] +
                        czn[k - 1 + 2 * nk] * (Tm1[I3(i, j, k)] * (1.0 - upmsk) + Tm1[I3(i, j, kp1)] * upmsk));
        }
        if (flux_z) flux_z[I3(i, j, k)] = g->dat[I2(i, j)] * ft2;
        result[C3(i, j, k)] = tmk[I3(i, j, k)] * (ft1[C2(i, j)] - ft2);
        ft1[C2(i, j)] = ft2;
      }
  }
  free(qx); free(qy); free(cxp); free(cxn); free(cyp); free(cyn);
  free(qz); free(czp); free(czn); free(ft1);
}

/* ------------------------------------------------------------------ */
/* ocean_tracer.F90:2788-2892 compute_tmask_limit                       */
/* ------------------------------------------------------------------ */
void orc_compute_tmask_limit(const orc_dom *d, const orc_grid *g, int ntr,
    const double *const *Tt, const double *mn, const double *mx,
    int index_temp, int index_salt, int ts_same, double *const *lim) {
  const int nk = d->nk;
  const double *tm = g->tmask;
  for (int n = 0; n < ntr; ++n) {
    const double *T = Tt[n];
    double *L = lim[n];
    PAR_FOR
    for (int k = 1; k <= nk; ++k)
      for (int j = d->jsc - 1; j <= d->jec; ++j)
        for (int i = d->isc - 1; i <= d->iec; ++i) {
          L[I3(i, j, k)] = 0.0;
          if (tm[I3(i, j, k)] == 1.0 && (T[I3(i, j, k)] < mn[n] || T[I3(i, j, k)] > mx[n]))
            L[I3(i, j, k)] = tm[I3(i, j, k)];
          if (tm[I3(i + 1, j, k)] == 1.0 && (T[I3(i + 1, j, k)] < mn[n] || T[I3(i + 1, j, k)] > mx[n]))
            L[I3(i, j, k)] = tm[I3(i, j, k)];
          if (tm[I3(i, j + 1, k)] == 1.0 && (T[I3(i, j + 1, k)] < mn[n] || T[I3(i, j + 1, k)] > mx[n]))
            L[I3(i, j, k)] = tm[I3(i, j, k)];
        }
    PAR_FOR
    for (int k = 1; k <= nk - 1; ++k)
      for (int j = d->jsc - 1; j <= d->jec; ++j)
        for (int i = d->isc - 1; i <= d->iec; ++i)
          if (tm[I3(i, j, k + 1)] == 1.0 && (T[I3(i, j, k + 1)] < mn[n] || T[I3(i, j, k + 1)] > mx[n]))
            L[I3(i, j, k)] = tm[I3(i, j, k)];
  }
  if (ts_same && index_temp >= 0 && index_salt >= 0) {
    double *Lt = lim[index_temp], *Ls = lim[index_salt];
    PAR_FOR
    for (int k = 1; k <= nk; ++k)
      for (int j = d->jsc - 1; j <= d->jec; ++j)
        for (int i = d->isc - 1; i <= d->iec; ++i)
          if (Lt[I3(i, j, k)] == 1.0 || Ls[I3(i, j, k)] == 1.0) {
            Lt[I3(i, j, k)] = 1.0;
            Ls[I3(i, j, k)] = 1.0;
          }
  }
}

/* ocean_tracer.F90:2341-2350 */
void orc_update_tracer_explicit(const orc_dom *d, const double *Tm1,
    const double *tend, const double *rdz_m1, const double *rdzr,
    double dtime, double *Tp1) {
  PAR_FOR
  for (int k = 1; k <= d->nk; ++k)
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        const size_t o = I3(i, j, k);
        Tp1[o] = (rdz_m1[o] * Tm1[o] + dtime * tend[o]) * rdzr[o];
      }
}

/* ------------------------------------------------------------------ */
/* ocean_util.F90:153-222 invtri                                        */
/* ------------------------------------------------------------------ */
void orc_invtri(const orc_dom *d, double *z, const double *topbc,
    const double *botbc, const double *dcb, double tdt, const int *kmz,
    const double *mask, const double *dh, const double *dhw, double aidif) {
  const int nk = d->nk;
  const int ni = NI;
  const double eps = 1.0e-30;
  const double tdt_aidif = tdt * aidif;
  /* work rows a,b,c,e,f(isd:ied,0:nk), bet(isd:ied): one set per thread
   * of the all-core build, rows j independent */
#pragma omp parallel
  {
  double *a = (double *)calloc((size_t)ni * (nk + 1), sizeof(double));
  double *b = (double *)calloc((size_t)ni * (nk + 1), sizeof(double));
  double *c = (double *)calloc((size_t)ni * (nk + 1), sizeof(double));
  double *e = (double *)calloc((size_t)ni * (nk + 1), sizeof(double));
  double *f = (double *)calloc((size_t)ni * (nk + 1), sizeof(double));
  double *bet = (double *)calloc((size_t)ni, sizeof(double));
#define W(arr, i, k) arr[(size_t)((i) - d->isd) + (size_t)ni * (size_t)(k)]
#pragma omp for schedule(static)
  for (int j = d->jsc; j <= d->jec; ++j) {
    for (int k = 1; k <= nk; ++k) {
      const int km1 = (k - 1 > 1) ? k - 1 : 1;
      const int kp1 = (k + 1 < nk) ? k + 1 : nk;
      for (int i = d->isc; i <= d->iec; ++i) {
        const double factu = tdt_aidif / (dhw[I3W(i, j, k - 1)] * dh[I3(i, j, k)]);
        const double factl = tdt_aidif / (dhw[I3W(i, j, k)] * dh[I3(i, j, k)]);
        W(a, i, k) = -dcb[I3(i, j, km1)] * factu * mask[I3(i, j, k)];
        W(c, i, k) = -dcb[I3(i, j, k)] * factl * mask[I3(i, j, kp1)];
        W(f, i, k) = z[I3(i, j, k)] * mask[I3(i, j, k)];
        W(b, i, k) = 1.0 - W(a, i, k) - W(c, i, k);
      }
    }
    for (int i = d->isc; i <= d->iec; ++i) {
      W(a, i, 1) = 0.0;
      W(c, i, nk) = 0.0;
      W(b, i, 1) = 1.0 - W(a, i, 1) - W(c, i, 1);
      W(b, i, nk) = 1.0 - W(a, i, nk) - W(c, i, nk);
      W(f, i, 1) = z[I3(i, j, 1)] + topbc[I2(i, j)] * tdt_aidif * mask[I3(i, j, 1)] / dh[I3(i, j, 1)];
      const int k = (kmz[I2(i, j)] > 2) ? kmz[I2(i, j)] : 2;
      W(f, i, k) = z[I3(i, j, k)] - botbc[I2(i, j)] * tdt_aidif * mask[I3(i, j, k)] / dh[I3(i, j, k)];
    }
    for (int i = d->isc; i <= d->iec; ++i) {
      bet[i - d->isd] = mask[I3(i, j, 1)] / (W(b, i, 1) + eps);
      z[I3(i, j, 1)] = W(f, i, 1) * bet[i - d->isd];
    }
    for (int k = 2; k <= nk; ++k)
      for (int i = d->isc; i <= d->iec; ++i) {
        W(e, i, k) = W(c, i, k - 1) * bet[i - d->isd];
        bet[i - d->isd] = mask[I3(i, j, k)] / (W(b, i, k) - W(a, i, k) * W(e, i, k) + eps);
        z[I3(i, j, k)] = (W(f, i, k) - W(a, i, k) * z[I3(i, j, k - 1)]) * bet[i - d->isd];
      }
    for (int k = nk - 1; k >= 1; --k)
      for (int i = d->isc; i <= d->iec; ++i)
        z[I3(i, j, k)] = z[I3(i, j, k)] - W(e, i, k + 1) * z[I3(i, j, k + 1)];
  }
#undef W
  free(a); free(b); free(c); free(e); free(f); free(bet);
  }
}

/* ocean_vert_mix.F90:3426-3513 (the part feeding invtri) */
void orc_vert_diffuse_implicit(const orc_dom *d, const orc_grid *g, int ntr,
    double *const *Tp1, const double *const *stf, const double *const *btf,
    const double *diff_cbt, const double *const *K33, int index_salt,
    double rho0, const double *rdz_p1, const double *dzwt, double dtime_t,
    double aidif) {
  const int nk = d->nk;
  const size_t n3 = P2 * (size_t)nk;
  if (aidif == 0.0) return;
  double *wrk2 = (double *)calloc(n3, sizeof(double));
  for (int n = 0; n < ntr; ++n) {
    const int nmix = (n == index_salt) ? 2 : 1;
    const double *dc = diff_cbt + n3 * (size_t)(nmix - 1);
    PAR_FOR
    for (int k = 1; k <= nk; ++k)
      for (int j = d->jsc; j <= d->jec; ++j)
        for (int i = d->isc; i <= d->iec; ++i) {
          const size_t o = I3(i, j, k);
          wrk2[o] = rho0 * dc[o] + ((K33 && K33[n]) ? K33[n][o] : 0.0);
        }
    orc_invtri(d, Tp1[n], stf[n], btf[n], wrk2, dtime_t, g->kmt, g->tmask, rdz_p1, dzwt, aidif);
  }
  free(wrk2);
}

/* ------------------------------------------------------------------ */
/* ocean_operators.F90:507-608 DIV_UD (B-grid, non-legacy ordering),    */
/* halo_in = halo_out = 1, i.e. on the data domain.                     */
/* ------------------------------------------------------------------ */
static void orc_div_ud(const orc_dom *d, const orc_grid *g, const double *ud,
                       double *out) {
  const double *u1 = ud, *u2 = ud + P2;
  for (size_t o = 0; o < P2; ++o) out[o] = 0.0;
  PAR_FOR
  for (int j = d->jsd + 1; j <= d->jed; ++j) {
    int i = d->isd + 1;
    double uhim = u1[I2(i - 1, j)] * g->dyu[I2(i - 1, j)];
    double vhim = u2[I2(i - 1, j)] * g->dxu[I2(i - 1, j)];
    double uhimjm = u1[I2(i - 1, j - 1)] * g->dyu[I2(i - 1, j - 1)];
    double vhimjm = u2[I2(i - 1, j - 1)] * g->dxu[I2(i - 1, j - 1)];
    double uhim_bay = 0.5 * (uhim + uhimjm);
    for (i = d->isd + 1; i <= d->ied; ++i) {
      const double uh = u1[I2(i, j)] * g->dyu[I2(i, j)];
      const double vh = u2[I2(i, j)] * g->dxu[I2(i, j)];
      const double uhjm = u1[I2(i, j - 1)] * g->dyu[I2(i, j - 1)];
      const double vhjm = u2[I2(i, j - 1)] * g->dxu[I2(i, j - 1)];
      const double uh_bay = 0.5 * (uh + uhjm);
      const double vh_bax = 0.5 * (vh + vhim);
      const double vhjm_bax = 0.5 * (vhjm + vhimjm);
      out[I2(i, j)] = ((uh_bay - uhim_bay) + (vh_bax - vhjm_bax)) * g->datr[I2(i, j)];
      uhim = uh; vhim = vh; uhimjm = uhjm; vhimjm = vhjm; uhim_bay = uh_bay;
    }
  }
}

/* GRAD_BAROTROPIC_P on the B-grid (ocean_operators.F90:631-676), defined
 * with the barotropic closure below; the subcycle loops use the same one. */
static void orc_grad_barotropic_p(const orc_dom *d, const orc_grid *g, const double *p,
                                  double *out);

/* ------------------------------------------------------------------ */
/* ocean_barotropic.F90:3225-3579 pred_corr_tropic_depth_bgrid with the  */
/* defaults barotropic_halo=1, use_legacy_barotropic_halos=.true., no    */
/* tides/geoid/friction/smoothing/OBC, and the tail :2804-2894.          */
/* ------------------------------------------------------------------ */
void orc_update_ocean_barotropic(const orc_dom *d, const orc_grid *g,
    const orc_bt_params *p, const double *eta_t_tau,
    const double *eta_t_bar_tau, const double *udrho_tau,
    const double *forcing_bt, const double *thicku, const double *patm,
    const double *pme, const double *river, const double *source,
    const double *eta_smooth, const double *rho_g, double *eta_bar,
    double *udrho, double *conv_rho_ud_t) {
  const int nts = p->nts;
  const double dtbt = p->dtbt;
  const double rho0r = 1.0 / p->rho0;
  const double dtbt_gamma_rho0r = dtbt * p->pred_corr_gamma * rho0r;  /* :1196 */
  const double dtbt_rho0r = dtbt * rho0r;                               /* :1197 */
  const double *tmask1 = g->tmask;   /* plane k=1 */
  double *eta_bt = (double *)calloc(2 * P2, sizeof(double));
  double *ud_bt = (double *)calloc(4 * P2, sizeof(double));    /* (…,2,2) */
  double *sf = (double *)malloc(P2 * sizeof(double));
  double *tmp = (double *)malloc(P2 * sizeof(double));
  double *ps = (double *)malloc(P2 * sizeof(double));
  double *gp = (double *)malloc(2 * P2 * sizeof(double));
  double *cori1 = (double *)malloc(P2 * sizeof(double));
  double *cori2 = (double *)malloc(P2 * sizeof(double));
#define ETA(o, lev) eta_bt[(o) + P2 * (size_t)((lev) - 1)]
#define UD(o, c, lev) ud_bt[(o) + P2 * ((size_t)((c) - 1) + 2 * (size_t)((lev) - 1))]
  for (size_t o = 0; o < P2; ++o) {          /* :1283-1288 */
    cori1[o] = 0.5 * dtbt * g->f[o];
    cori2[o] = 1.0 / (1.0 + cori1[o] * cori1[o]);
  }
  int itime = 1;
  int fstau = (itime + 0) % 2 + 1;
  for (size_t o = 0; o < P2; ++o) {          /* :3281-3318 */
    eta_bar[o] = p->initsum_with_bar ? eta_t_bar_tau[o] : eta_t_tau[o];
    udrho[o] = 0.0;
    udrho[o + P2] = 0.0;
    ETA(o, fstau) = eta_t_bar_tau[o];
    UD(o, 1, fstau) = udrho_tau[o];
    UD(o, 2, fstau) = udrho_tau[o + P2];
    sf[o] = tmask1[o] * (pme[o] + river[o] + source[o] - eta_smooth[o]);
  }
  for (itime = 1; itime <= nts; ++itime) {
    fstau = (itime + 0) % 2 + 1;
    const int fstaup1 = (itime + 1) % 2 + 1;
    /* predictor (:3365-3376) */
    double *ubuf = &UD(0, 1, fstau);
    orc_div_ud(d, g, ubuf, tmp);
    PAR_FOR
    for (size_t o = 0; o < P2; ++o) tmp[o] = -tmp[o] * tmask1[o];
    PAR_FOR
    for (size_t o = 0; o < P2; ++o)
      ETA(o, fstaup1) = ETA(o, fstau) + dtbt_gamma_rho0r * (tmp[o] + sf[o]);
    /* legacy halo update of eta_t_bt(fstaup1) (:3383) */
    orc_halo_update(d, 1, &ETA(0, fstaup1), 1, 3);
    PAR_FOR
    for (size_t o = 0; o < P2; ++o) ps[o] = rho_g[o] * ETA(o, fstaup1) + patm[o];
    orc_grad_barotropic_p(d, g, ps, gp);
    PAR_FOR
    for (size_t o = 0; o < P2; ++o) {       /* :3464-3486 */
      const double pf1 = -(thicku[o] * gp[o]);
      const double pf2 = -(thicku[o] * gp[o + P2]);
      const double u1 = UD(o, 1, fstau), u2 = UD(o, 2, fstau);
      const double t1 = u1 + dtbt * (0.5 * g->f[o] * u2 + pf1 + forcing_bt[o] + 0.0);
      const double t2 = u2 + dtbt * (-(0.5 * g->f[o] * u1) + pf2 + forcing_bt[o + P2] + 0.0);
      UD(o, 1, fstaup1) = cori2[o] * (t1 + cori1[o] * t2);
      UD(o, 2, fstaup1) = cori2[o] * (t2 - cori1[o] * t1);
    }
    /* corrector (:3497-3505) */
    orc_div_ud(d, g, &UD(0, 1, fstaup1), tmp);
    PAR_FOR
    for (size_t o = 0; o < P2; ++o) tmp[o] = -tmp[o] * tmask1[o];
    PAR_FOR
    for (size_t o = 0; o < P2; ++o)
      ETA(o, fstaup1) = ETA(o, fstau) + dtbt_rho0r * (tmp[o] + sf[o]);
    /* accumulate (:3532-3549) */
    const double wgt = p->stepping_a ? 1.0 : (double)(nts - itime + 1);
    PAR_FOR
    for (size_t o = 0; o < P2; ++o) {
      eta_bar[o] = eta_bar[o] + ETA(o, fstaup1);
      if (p->stepping_a) {
        udrho[o] = udrho[o] + UD(o, 1, fstaup1);
        udrho[o + P2] = udrho[o + P2] + UD(o, 2, fstaup1);
      } else {
        udrho[o] = udrho[o] + wgt * UD(o, 1, fstaup1);
        udrho[o + P2] = udrho[o + P2] + wgt * UD(o, 2, fstaup1);
      }
    }
  }
  /* tail of update_ocean_barotropic (:2804-2811, :2846-2894) */
  const double rntsp1 = 1.0 / ((double)nts + 1.0);
  const double rnts = p->stepping_a ? 1.0 / (double)nts
                                    : 2.0 / ((double)nts * (double)(nts + 1));
  for (size_t o = 0; o < P2; ++o) eta_bar[o] = rntsp1 * eta_bar[o];
  orc_halo_update(d, 1, eta_bar, 1, 3);
  for (size_t o = 0; o < 2 * P2; ++o) udrho[o] = rnts * udrho[o];
  orc_halo_update(d, 1, udrho, 2, 3);
  if (conv_rho_ud_t) {
    orc_div_ud(d, g, udrho, conv_rho_ud_t);
    PAR_FOR
    for (size_t o = 0; o < P2; ++o) conv_rho_ud_t[o] = -conv_rho_ud_t[o];
    orc_halo_update(d, 1, conv_rho_ud_t, 1, 3);
  }
#undef ETA
#undef UD
  free(eta_bt); free(ud_bt); free(sf); free(tmp); free(ps); free(gp);
  free(cori1); free(cori2);
}

/* ocean_tracer_diag.F90:2648-2661 */
double orc_tracer_content(const orc_dom *d, const orc_grid *g,
    const double *T, const double *rho_dzt) {
  long double tot = 0.0L;
  for (int j = d->jsc; j <= d->jec; ++j)
    for (int i = d->isc; i <= d->iec; ++i) {
      long double col = 0.0L;
      for (int k = 1; k <= d->nk; ++k) {
        const size_t o = I3(i, j, k);
        col += (long double)T[o] * (long double)(rho_dzt[o] * g->dat[I2(i, j)] * g->tmask[o]);
      }
      tot += col;
    }
  return (double)tot;
}

/* ---- ocean_advection_velocity, B-grid branch ----------------------------
 * ocean_core/ocean_advection_velocity.F90:568-672 (no blobs, OBC or inflow):
 *   uhrho_et = BAY(u(:,:,k,1)*dyu*rho_dzu(k))/dyte              (:601-602)
 *   vhrho_nt = BAX(u(:,:,k,2)*dxu*rho_dzu(k))/dxtn              (:603-604)
 *   diverge_t = tmask*(BDX_ET(uhrho_et) + BDY_NT(vhrho_nt))     (:607-608)
 *   wrho_bt(0) = -(pme + river)                                 (:615)
 *   wrho_bt(k) = (tmp + diverge_t + wrho_bt(k-1))*tmask,
 *   tmp = rho_dzt_tendency - mass_source                        (:633-636)
 * with the operators of ocean_core/ocean_operators.F90 (halo = 1):
 *   BAX :834-846, BAY :889-902, BDX_ET :945-958, BDY_NT :1230-1243.
 * Each operator is evaluated over the whole data domain into its own array,
 * as the reference's array-valued functions are. */
static void orc_bay(const orc_dom *d, const double *a, double *out) {
  for (int j = d->jsd + 1; j <= d->jed; ++j)
    for (int i = d->isd; i <= d->ied; ++i) out[I2(i, j)] = 0.5 * (a[I2(i, j)] + a[I2(i, j - 1)]);
  for (int i = d->isd; i <= d->ied; ++i) out[I2(i, d->jsd)] = 0.0;
}

static void orc_bax(const orc_dom *d, const double *a, double *out) {
  for (int j = d->jsd; j <= d->jed; ++j) {
    for (int i = d->isd + 1; i <= d->ied; ++i)
      out[I2(i, j)] = 0.5 * (a[I2(i, j)] + a[I2(i - 1, j)]);
    out[I2(d->isd, j)] = 0.0;
  }
}

void orc_advection_velocity(const orc_dom *d, const orc_grid *g, const double *u,
                            const double *rho_dzu, const double *pme, const double *river,
                            const double *rho_dzt_tendency, const double *mass_source,
                            double *uhrho_et, double *vhrho_nt, double *wrho_bt,
                            double *diverge_t) {
  const size_t p2 = P2, p3 = p2 * (size_t)d->nk;
  double *a = (double *)malloc(sizeof(double) * p2);
  double *b = (double *)malloc(sizeof(double) * p2);
  double *bdx = (double *)malloc(sizeof(double) * p2);
  double *bdy = (double *)malloc(sizeof(double) * p2);
  double *div = (double *)malloc(sizeof(double) * p3);
  for (int k = 1; k <= d->nk; ++k) {
    for (int j = d->jsd; j <= d->jed; ++j)
      for (int i = d->isd; i <= d->ied; ++i)
        a[I2(i, j)] = u[I3(i, j, k)] * g->dyu[I2(i, j)] * rho_dzu[I3(i, j, k)];
    orc_bay(d, a, b);
    for (int j = d->jsd; j <= d->jed; ++j)
      for (int i = d->isd; i <= d->ied; ++i) uhrho_et[I3(i, j, k)] = b[I2(i, j)] / g->dyte[I2(i, j)];
    for (int j = d->jsd; j <= d->jed; ++j)
      for (int i = d->isd; i <= d->ied; ++i)
        a[I2(i, j)] = u[p3 + I3(i, j, k)] * g->dxu[I2(i, j)] * rho_dzu[I3(i, j, k)];
    orc_bax(d, a, b);
    for (int j = d->jsd; j <= d->jed; ++j)
      for (int i = d->isd; i <= d->ied; ++i) vhrho_nt[I3(i, j, k)] = b[I2(i, j)] / g->dxtn[I2(i, j)];
    /* BDX_ET(uhrho_et) and BDY_NT(vhrho_nt) */
    for (int j = d->jsd; j <= d->jed; ++j) {
      for (int i = d->isd + 1; i <= d->ied; ++i)
        bdx[I2(i, j)] = (g->dyte[I2(i, j)] * uhrho_et[I3(i, j, k)] -
                         g->dyte[I2(i - 1, j)] * uhrho_et[I3(i - 1, j, k)]) * g->datr[I2(i, j)];
      bdx[I2(d->isd, j)] = 0.0;
    }
    for (int j = d->jsd + 1; j <= d->jed; ++j)
      for (int i = d->isd; i <= d->ied; ++i)
        bdy[I2(i, j)] = (g->dxtn[I2(i, j)] * vhrho_nt[I3(i, j, k)] -
                         g->dxtn[I2(i, j - 1)] * vhrho_nt[I3(i, j - 1, k)]) * g->datr[I2(i, j)];
    for (int i = d->isd; i <= d->ied; ++i) bdy[I2(i, d->jsd)] = 0.0;
    for (int j = d->jsd; j <= d->jed; ++j)
      for (int i = d->isd; i <= d->ied; ++i)
        div[I3(i, j, k)] = g->tmask[I3(i, j, k)] * (bdx[I2(i, j)] + bdy[I2(i, j)]);
  }
  for (int j = d->jsd; j <= d->jed; ++j)
    for (int i = d->isd; i <= d->ied; ++i) wrho_bt[I3W(i, j, 0)] = -(pme[I2(i, j)] + river[I2(i, j)]);
  for (int k = 1; k <= d->nk; ++k)
    for (int j = d->jsd; j <= d->jed; ++j)
      for (int i = d->isd; i <= d->ied; ++i) {
        const double rt = rho_dzt_tendency ? rho_dzt_tendency[I3(i, j, k)] : 0.0;
        const double ms = mass_source ? mass_source[I3(i, j, k)] : 0.0;
        const double tmp = rt - ms;
        wrho_bt[I3W(i, j, k)] = (tmp + div[I3(i, j, k)] + wrho_bt[I3W(i, j, k - 1)]) *
                                g->tmask[I3(i, j, k)];
      }
  if (diverge_t) memcpy(diverge_t, div, sizeof(double) * p3);
  free(a);
  free(b);
  free(bdx);
  free(bdy);
  free(div);
}

/* ---- ocean_barotropic_forcing ---------------------------------------------
 * ocean_core/ocean_barotropic.F90:3058-3141 with the default namelist
 * (zero_nonlinear_forcing_bt = zero_forcing_bt = udrho_lap = udrho_bih =
 * .false., no OBC).  On the B-grid tmasken(:,:,:,n) = umask
 * (ocean_core/ocean_grids.F90:1403-1410):
 *   forcing_bt = 0 on the data domain;
 *   wrk(n) = sum_k umask(k)*accel(k,n) over the compute domain (k ascending);
 *   forcing_bt(n) = wrk(n) + umask(1)*(smf(n) - bmf(n));
 *   mpp_update_domains (single PE: cyclic wrap);
 *   forcing_bt(n) = forcing_bt(n) + friction_lap(n) + friction_bih(n) (both 0).
 * accel: (isd:ied,jsd:jed,nk,2); smf, bmf, forcing_bt: (isd:ied,jsd:jed,2). */
void orc_barotropic_forcing(const orc_dom *d, const orc_grid *g, const double *accel,
                            const double *smf, const double *bmf, double *forcing_bt) {
  const size_t p2 = P2, p3 = p2 * (size_t)d->nk;
  for (size_t o = 0; o < 2 * p2; ++o) forcing_bt[o] = 0.0;
  for (int n = 0; n < 2; ++n)
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        double wrk = 0.0;
        for (int k = 1; k <= d->nk; ++k)
          wrk = wrk + g->umask[I3(i, j, k)] * accel[p3 * (size_t)n + I3(i, j, k)];
        forcing_bt[p2 * (size_t)n + I2(i, j)] =
            wrk + g->umask[I3(i, j, 1)] * (smf[p2 * (size_t)n + I2(i, j)] -
                                           bmf[p2 * (size_t)n + I2(i, j)]);
      }
  orc_halo_update(d, 1, forcing_bt, 2, 3);
  for (size_t o = 0; o < 2 * p2; ++o) forcing_bt[o] = forcing_bt[o] + 0.0 + 0.0;
}

/* ocean_barotropic.F90:2667-2700 (eta_and_pbot_tendency, use_blobs false)
 * followed by :2290-2363 (eta_and_pbot_update, no OBC, truncate_eta false,
 * splitting).  Depth-based (pressure_based = 0):
 *   tendency = tmask(1)*rho0r*(conv + pme + river + source) on the data domain;
 *   x_taup1  = x_taum1 + dtime*tendency on the compute domain, halo-updated.
 * Pressure-based: tendency = dpatm_dt + grav*tmask(1)*(conv + ...), x is
 * anompb, and pbot_taup1 = x_taup1 + rho0grav*ht on the data domain. */
void orc_eta_and_pbot_update(const orc_dom *d, const orc_grid *g, int pressure_based,
    double rho0, double grav, double dtime, const double *conv_tau, const double *pme,
    const double *river, const double *source, const double *dpatm_dt, const double *ht,
    const double *x_taum1, double *x_taup1, double *pbot_taup1, double *tendency) {
  const double rho0r = 1.0 / rho0;          /* ocean_parameters.F90:260 */
  const double rho0grav = rho0 * grav;      /* ocean_barotropic.F90:1191 */
  for (int j = d->jsd; j <= d->jed; ++j)
    for (int i = d->isd; i <= d->ied; ++i) {
      const size_t o = I2(i, j);
      const double tm = g->tmask[I3(i, j, 1)];
      if (!pressure_based)
        tendency[o] = tm * rho0r * (conv_tau[o] + pme[o] + river[o] + source[o]);
      else
        tendency[o] = dpatm_dt[o] + grav * tm * (conv_tau[o] + pme[o] + river[o] + source[o]);
    }
  for (int j = d->jsc; j <= d->jec; ++j)
    for (int i = d->isc; i <= d->iec; ++i)
      x_taup1[I2(i, j)] = x_taum1[I2(i, j)] + dtime * tendency[I2(i, j)];
  orc_halo_update(d, 1, x_taup1, 1, 3);
  if (pressure_based)
    for (size_t o = 0; o < P2; ++o) pbot_taup1[o] = x_taup1[o] + rho0grav * ht[o];
}

/* ocean_operators.F90:462-477, REMAP_BT_TO_BU with halo 1. */
static void orc_remap_bt_to_bu(const orc_dom *d, const double *a, const double *dte,
    const double *dtw, const double *dus, const double *dun, const double *daur, double *out) {
  for (size_t o = 0; o < P2; ++o) out[o] = 0.0;
  for (int j = d->jsd; j <= d->jed - 1; ++j)
    for (int i = d->isd; i <= d->ied - 1; ++i)
      out[I2(i, j)] = (a[I2(i, j)] * dte[I2(i, j)] * dus[I2(i, j)]
                       + a[I2(i + 1, j)] * dtw[I2(i + 1, j)] * dus[I2(i, j)]
                       + a[I2(i, j + 1)] * dte[I2(i, j + 1)] * dun[I2(i, j)]
                       + a[I2(i + 1, j + 1)] * dtw[I2(i + 1, j + 1)] * dun[I2(i, j)])
                      * daur[I2(i, j)];
}

/* ocean_operators.F90:631-676, GRAD_BAROTROPIC_P on the B-grid with
 * halo_in = halo_out = 1; umask_bt = umask(:,:,1), dxur_bt = dxur
 * (:321-323).  out: (isd:ied, jsd:jed, 2). */
static void orc_grad_barotropic_p(const orc_dom *d, const orc_grid *g, const double *p,
                                  double *out) {
  for (size_t o = 0; o < 2 * P2; ++o) out[o] = 0.0;
  PAR_FOR
  for (int j = d->jsd; j <= d->jed - 1; ++j)
    for (int i = d->isd; i <= d->ied - 1; ++i) {
      const double p_fay = 0.5 * (p[I2(i, j)] + p[I2(i, j + 1)]);
      const double p_fax = 0.5 * (p[I2(i, j)] + p[I2(i + 1, j)]);
      const double pip_fay = 0.5 * (p[I2(i + 1, j)] + p[I2(i + 1, j + 1)]);
      const double pjp_fax = 0.5 * (p[I2(i, j + 1)] + p[I2(i + 1, j + 1)]);
      out[I2(i, j)] = (pip_fay - p_fay) * g->dxur[I2(i, j)] * g->umask[I3(i, j, 1)];
      out[P2 + I2(i, j)] = (pjp_fax - p_fax) * g->dyur[I2(i, j)] * g->umask[I3(i, j, 1)];
    }
}

/* ocean_barotropic.F90:2906-2921 and :2956-2985 (splitting, no OBC, no
 * tidal or geoid forcing, zero_eta_u false):
 *   depth-based:    eta_u = umask(1)*REMAP_BT_TO_BU(eta_t_bar), halo-updated;
 *   pressure-based: eta_u (pbot_u) = umask(1)*REMAP_BT_TO_BU(rho0grav*ht + anompb_bar);
 *   ps = rho_g*eta_t_bar + patm; grad_ps = GRAD_BAROTROPIC_P(ps);
 *   grad_anompb = GRAD_BAROTROPIC_P(anompb_bar).
 * eta_t_bar and anompb_bar are the itime level (taup1 for two-level
 * stepping, :2797-2801).  ht is read only when pressure_based;
 * anompb_bar may be NULL when depth-based (grad_anompb is then not written).
 * tmasken (Grd%tmasken(:,:,1,1:2)) selects the C-grid gradient
 * (ocean_operators.F90:664-669); NULL: B-grid. */
static void orc_grad_p_c(const orc_dom *d, const double *p, const double *dxter,
                         const double *dytnr, const double *tmasken, double *out);

void orc_barotropic_closure(const orc_dom *d, const orc_grid *g, int pressure_based,
    double rho0, double grav, const double *eta_t_bar, const double *anompb_bar,
    const double *patm, const double *rho_g, const double *ht, const double *tmasken,
    const double *dte, const double *dtw, const double *dus, const double *dun,
    const double *daur, double *eta_u, double *ps, double *grad_ps, double *grad_anompb) {
  const double rho0grav = rho0 * grav;
  double *a = (double *)malloc(P2 * sizeof(double));
  double *r = (double *)malloc(P2 * sizeof(double));
  if (!pressure_based) {
    for (size_t o = 0; o < P2; ++o) a[o] = eta_t_bar[o];
  } else {
    for (size_t o = 0; o < P2; ++o) a[o] = rho0grav * ht[o] + anompb_bar[o];
  }
  orc_remap_bt_to_bu(d, a, dte, dtw, dus, dun, daur, r);
  for (int j = d->jsd; j <= d->jed; ++j)
    for (int i = d->isd; i <= d->ied; ++i)
      eta_u[I2(i, j)] = g->umask[I3(i, j, 1)] * r[I2(i, j)];
  orc_halo_update(d, 1, eta_u, 1, 3);
  for (size_t o = 0; o < P2; ++o) ps[o] = rho_g[o] * eta_t_bar[o] + patm[o];
  if (!tmasken) {
    orc_grad_barotropic_p(d, g, ps, grad_ps);
    if (anompb_bar && grad_anompb) orc_grad_barotropic_p(d, g, anompb_bar, grad_anompb);
  } else {
    for (size_t o = 0; o < P2; ++o) {
      a[o] = 1.0 / (g->dxte[o] + 1.0e-40);   /* Grd%dxter, ocean_grids.F90:956 */
      r[o] = 1.0 / (g->dytn[o] + 1.0e-40);   /* Grd%dytnr, :959 */
    }
    orc_grad_p_c(d, ps, a, r, tmasken, grad_ps);
    if (anompb_bar && grad_anompb) orc_grad_p_c(d, anompb_bar, a, r, tmasken, grad_anompb);
  }
  free(a);
  free(r);
}

/* ------------------------------------------------------------------ */
/* The four split-explicit barotropic schemes of update_ocean_barotropic */
/* (ocean_barotropic.F90:2829-2843) with barotropic_halo = 1, no tides,  */
/* geoid, friction, smoothing, blobs or OBC:                            */
/*   pred_corr_tropic_depth_bgrid :3225-3579 (legacy halos),            */
/*   pred_corr_tropic_depth_cgrid :3601-3965,                           */
/*   pred_corr_tropic_press_bgrid :3989-4329 (legacy halos),            */
/*   pred_corr_tropic_press_cgrid :4355-4693,                           */
/* and the tail :2846-2894 (time means, conv_rho_ud_t).                 */
/* ------------------------------------------------------------------ */

/* ocean_operators.F90:526-536, DIV_UD on the C-grid, halo 1.
 * dxter = 1/(dxte+epsln), dytnr = 1/(dytn+epsln) (ocean_grids.F90:956,959). */
static void orc_div_ud_c(const orc_dom *d, const orc_grid *g, const double *ud,
                         const double *dxter, const double *dytnr, double *out) {
  const double *u1 = ud, *u2 = ud + P2;
  for (size_t o = 0; o < P2; ++o) out[o] = 0.0;
  for (int j = d->jsd + 1; j <= d->jed; ++j)
    for (int i = d->isd + 1; i <= d->ied; ++i) {
      const double uhim = u1[I2(i - 1, j)] * g->dat[I2(i - 1, j)] * dxter[I2(i - 1, j)];
      const double vhjm = u2[I2(i, j - 1)] * g->dat[I2(i, j - 1)] * dytnr[I2(i, j - 1)];
      const double uh = u1[I2(i, j)] * g->dat[I2(i, j)] * dxter[I2(i, j)];
      const double vh = u2[I2(i, j)] * g->dat[I2(i, j)] * dytnr[I2(i, j)];
      out[I2(i, j)] = ((uh - uhim) + (vh - vhjm)) * g->datr[I2(i, j)];
    }
}

/* ocean_operators.F90:664-669, GRAD_BAROTROPIC_P on the C-grid, halo 1. */
static void orc_grad_p_c(const orc_dom *d, const double *p, const double *dxter,
                         const double *dytnr, const double *tmasken, double *out) {
  for (size_t o = 0; o < 2 * P2; ++o) out[o] = 0.0;
  for (int j = d->jsd; j <= d->jed - 1; ++j)
    for (int i = d->isd; i <= d->ied - 1; ++i) {
      out[I2(i, j)] = (p[I2(i + 1, j)] - p[I2(i, j)]) * dxter[I2(i, j)] * tmasken[I2(i, j)];
      out[P2 + I2(i, j)] =
          (p[I2(i, j + 1)] - p[I2(i, j)]) * dytnr[I2(i, j)] * tmasken[P2 + I2(i, j)];
    }
}

void orc_update_ocean_barotropic_variant(const orc_dom *d, const orc_grid *g,
    const orc_bt_params *p, int cgrid, int press, double grav, const double *x_tau,
    const double *xbar_tau, const double *udrho_tau, const double *forcing_bt,
    const double *thick, const double *tmasken, const double *patm, const double *rho_g,
    const double *dpatm_dt, const double *pme, const double *river, const double *source,
    const double *smooth, double *xbar, double *udrho, double *conv_rho_ud_t) {
  const int nts = p->nts;
  const double dtbt = p->dtbt, gamma = p->pred_corr_gamma;
  const double rho0r = 1.0 / p->rho0;
  const double dtbt_gamma = dtbt * gamma;                              /* :1195 */
  const double dtbt_gamma_rho0r = dtbt * gamma * rho0r;                /* :1196 */
  const double dtbt_rho0r = dtbt * rho0r;                              /* :1197 */
  const double ab0 = 23.0 / 12.0, ab1 = -16.0 / 12.0, ab2 = 5.0 / 12.0; /* :617-619 */
  const double *tmask1 = g->tmask;
  double *x_bt = (double *)calloc(2 * P2, sizeof(double));
  double *ud_bt = (double *)calloc(4 * P2, sizeof(double));
  double *cf = (double *)calloc(6 * P2, sizeof(double));    /* coriolis_force_bt(…,2,3) */
  double *acc = (double *)calloc(2 * P2, sizeof(double));   /* coriolis_accel_bt */
  double *sf = (double *)malloc(P2 * sizeof(double));
  double *tmp = (double *)malloc(P2 * sizeof(double));
  double *ps = (double *)malloc(P2 * sizeof(double));
  double *gp = (double *)malloc(2 * P2 * sizeof(double));
  double *cori1 = (double *)malloc(P2 * sizeof(double));
  double *cori2 = (double *)malloc(P2 * sizeof(double));
  double *dxter = (double *)malloc(P2 * sizeof(double));
  double *dytnr = (double *)malloc(P2 * sizeof(double));
#define XB(o, lev) x_bt[(o) + P2 * (size_t)((lev) - 1)]
#define UD(o, c, lev) ud_bt[(o) + P2 * ((size_t)((c) - 1) + 2 * (size_t)((lev) - 1))]
#define CF(o, c, lev) cf[(o) + P2 * ((size_t)((c) - 1) + 2 * (size_t)((lev) - 1))]
  for (size_t o = 0; o < P2; ++o) {
    cori1[o] = 0.5 * dtbt * g->f[o];                                  /* :1283-1288 */
    cori2[o] = 1.0 / (1.0 + cori1[o] * cori1[o]);
    dxter[o] = 1.0 / (g->dxte[o] + 1.0e-40);
    dytnr[o] = 1.0 / (g->dytn[o] + 1.0e-40);
  }
  int itime = 1;
  int fstau = (itime + 0) % 2 + 1, fstaup1 = (itime + 1) % 2 + 1;
  int m2 = (itime + 0) % 3 + 1, m1 = (itime + 1) % 3 + 1, m0 = (itime + 2) % 3 + 1;
  for (size_t o = 0; o < P2; ++o) {
    xbar[o] = p->initsum_with_bar ? xbar_tau[o] : x_tau[o];
    udrho[o] = 0.0;
    udrho[o + P2] = 0.0;
    XB(o, fstau) = xbar_tau[o];
    XB(o, fstaup1) = xbar_tau[o];     /* press: :4103; depth: overwritten */
    UD(o, 1, fstau) = udrho_tau[o];
    UD(o, 2, fstau) = udrho_tau[o + P2];
    if (!press)
      sf[o] = tmask1[o] * (pme[o] + river[o] + source[o] - smooth[o]);
    else
      sf[o] = tmask1[o] * (grav * (pme[o] + river[o] + source[o] - smooth[o]) + dpatm_dt[o]);
  }
  /* C-grid: Coriolis force at all three Adams-Bashforth levels (:3711-3730) */
#define CORI_FORCE(i, j, lev)                                                          \
  do {                                                                                 \
    CF(I2(i, j), 1, m0) = 0.25 * (g->f[I2(i, j)] * (UD(I2(i, j), 2, lev) +             \
                                                    UD(I2((i) + 1, j), 2, lev)) +      \
                                  g->f[I2(i, (j) - 1)] * (UD(I2(i, (j) - 1), 2, lev) + \
                                                          UD(I2((i) + 1, (j) - 1), 2, lev))); \
    CF(I2(i, j), 2, m0) = -(0.25 * (g->f[I2(i, j)] * (UD(I2(i, j), 1, lev) +           \
                                                      UD(I2(i, (j) + 1), 1, lev)) +    \
                                    g->f[I2((i) - 1, j)] * (UD(I2((i) - 1, j), 1, lev) + \
                                                            UD(I2((i) - 1, (j) + 1), 1, lev)))); \
  } while (0)
  if (cgrid)
    for (int j = d->jsc; j <= d->jec; ++j)
      for (int i = d->isc; i <= d->iec; ++i) {
        CORI_FORCE(i, j, fstau);
        for (int c = 1; c <= 2; ++c) {
          CF(I2(i, j), c, m1) = CF(I2(i, j), c, m0);
          CF(I2(i, j), c, m2) = CF(I2(i, j), c, m0);
        }
      }
  const int nthick = cgrid ? 2 : 1;
  for (itime = 1; itime <= nts; ++itime) {
    fstau = (itime + 0) % 2 + 1;
    fstaup1 = (itime + 1) % 2 + 1;
    m2 = (itime + 0) % 3 + 1;
    m1 = (itime + 1) % 3 + 1;
    m0 = (itime + 2) % 3 + 1;
    /* predictor */
    if ((press && cgrid) || gamma > 0.0) {   /* press_cgrid has no gamma test */
      if (cgrid) orc_div_ud_c(d, g, &UD(0, 1, fstau), dxter, dytnr, tmp);
      else orc_div_ud(d, g, &UD(0, 1, fstau), tmp);
      for (size_t o = 0; o < P2; ++o) {
        if (!press) {
          tmp[o] = -tmp[o] * tmask1[o];
          XB(o, fstaup1) = XB(o, fstau) + dtbt_gamma_rho0r * (tmp[o] + sf[o]);
        } else {
          tmp[o] = -grav * tmp[o];
          XB(o, fstaup1) = XB(o, fstau) + dtbt_gamma * (tmp[o] + sf[o]);
        }
      }
    } else {
      for (size_t o = 0; o < P2; ++o) XB(o, fstaup1) = XB(o, fstau);
    }
    if (!cgrid) orc_halo_update(d, 1, &XB(0, fstaup1), 1, 3);   /* legacy halos */
    for (size_t o = 0; o < P2; ++o)
      ps[o] = press ? tmask1[o] * XB(o, fstaup1) : rho_g[o] * XB(o, fstaup1) + patm[o];
    if (cgrid) {
      for (size_t o = 0; o < 2 * P2; ++o) acc[o] = 0.0;
      for (int j = d->jsd + 1; j <= d->jed - 1; ++j)
        for (int i = d->isd + 1; i <= d->ied - 1; ++i) {
          CORI_FORCE(i, j, fstau);
          for (int c = 1; c <= 2; ++c)
            acc[P2 * (size_t)(c - 1) + I2(i, j)] = ab0 * CF(I2(i, j), c, m0) +
                                                   ab1 * CF(I2(i, j), c, m1) +
                                                   ab2 * CF(I2(i, j), c, m2);
        }
      orc_grad_p_c(d, ps, dxter, dytnr, tmasken, gp);
    } else {
      orc_grad_barotropic_p(d, g, ps, gp);
    }
    for (size_t o = 0; o < P2; ++o) {
      double pf[2];
      for (int c = 0; c < 2; ++c) {
        const double th = thick[(nthick == 2 ? P2 * (size_t)c : 0) + o];
        pf[c] = press ? -(rho0r * th * gp[P2 * (size_t)c + o]) : -(th * gp[P2 * (size_t)c + o]);
      }
      const double u1 = UD(o, 1, fstau), u2 = UD(o, 2, fstau);
      if (cgrid) {
        const double a1 = acc[o] + pf[0] + forcing_bt[o];
        const double a2 = acc[o + P2] + pf[1] + forcing_bt[o + P2];
        UD(o, 1, fstaup1) = u1 + dtbt * tmasken[o] * (press ? a1 : a1 + 0.0);
        UD(o, 2, fstaup1) = u2 + dtbt * tmasken[o + P2] * (press ? a2 : a2 + 0.0);
      } else {
        double t1, t2;
        if (!press) {
          t1 = u1 + dtbt * (0.5 * g->f[o] * u2 + pf[0] + forcing_bt[o] + 0.0);
          t2 = u2 + dtbt * (-(0.5 * g->f[o] * u1) + pf[1] + forcing_bt[o + P2] + 0.0);
        } else {
          t1 = u1 + dtbt * (0.5 * g->f[o] * u2 + pf[0] + forcing_bt[o]);
          t2 = u2 + dtbt * (-(0.5 * g->f[o] * u1) + pf[1] + forcing_bt[o + P2]);
        }
        UD(o, 1, fstaup1) = cori2[o] * (t1 + cori1[o] * t2);
        UD(o, 2, fstaup1) = cori2[o] * (t2 - cori1[o] * t1);
      }
    }
    if (cgrid) orc_halo_update(d, 1, &UD(0, 1, fstaup1), 2, 3);
    /* corrector */
    if (cgrid) orc_div_ud_c(d, g, &UD(0, 1, fstaup1), dxter, dytnr, tmp);
    else orc_div_ud(d, g, &UD(0, 1, fstaup1), tmp);
    for (size_t o = 0; o < P2; ++o) {
      if (!press) {
        tmp[o] = -tmp[o] * tmask1[o];
        XB(o, fstaup1) = XB(o, fstau) + dtbt_rho0r * (tmp[o] + sf[o]);
      } else {
        tmp[o] = -grav * tmp[o];
        XB(o, fstaup1) = XB(o, fstau) + dtbt * (tmp[o] + sf[o]);
      }
    }
    /* accumulate */
    const double wgt = (double)(nts - itime + 1);
    for (size_t o = 0; o < P2; ++o) {
      xbar[o] = xbar[o] + XB(o, fstaup1);
      if (p->stepping_a) {
        udrho[o] = udrho[o] + UD(o, 1, fstaup1);
        udrho[o + P2] = udrho[o + P2] + UD(o, 2, fstaup1);
      } else {
        udrho[o] = udrho[o] + wgt * UD(o, 1, fstaup1);
        udrho[o + P2] = udrho[o + P2] + wgt * UD(o, 2, fstaup1);
      }
    }
  }
#undef CORI_FORCE
  const double rntsp1 = 1.0 / ((double)nts + 1.0);
  const double rnts = p->stepping_a ? 1.0 / (double)nts
                                    : 2.0 / ((double)nts * (double)(nts + 1));
  for (size_t o = 0; o < P2; ++o) xbar[o] = rntsp1 * xbar[o];
  orc_halo_update(d, 1, xbar, 1, 3);
  for (size_t o = 0; o < 2 * P2; ++o) udrho[o] = rnts * udrho[o];
  orc_halo_update(d, 1, udrho, 2, 3);
  if (conv_rho_ud_t) {
    if (cgrid) orc_div_ud_c(d, g, udrho, dxter, dytnr, conv_rho_ud_t);
    else orc_div_ud(d, g, udrho, conv_rho_ud_t);
    for (size_t o = 0; o < P2; ++o) conv_rho_ud_t[o] = -conv_rho_ud_t[o];
    orc_halo_update(d, 1, conv_rho_ud_t, 1, 3);
  }
#undef XB
#undef UD
#undef CF
  free(x_bt); free(ud_bt); free(cf); free(acc); free(sf); free(tmp); free(ps); free(gp);
  free(cori1); free(cori2); free(dxter); free(dytnr);
}
