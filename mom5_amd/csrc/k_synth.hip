// Seeded synthetic ocean state built on the device (include/mom5_synth.h).
//
// Benchmark plumbing, not reference routines: at 3600 x 2700 x 75 the 3-D
// inputs (5.8 GB per FP64 field) are generated where they are used.  Every
// kernel writes whole data-domain planes and reads only in-plane neighbours
// that exist (index checks on every edge).
#include "mom_internal.h"
#include "mom5_synth.h"

namespace {

constexpr int SB = 256;

__device__ inline unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// uniform in (0,1]
__device__ inline double unif(unsigned long long seed, unsigned long long idx, int s) {
  const unsigned long long h = splitmix64(seed ^ splitmix64(idx * 4ULL + (unsigned long long)s));
  return ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

// standard normal (Box-Muller on two hashed uniforms)
__device__ inline double normal(unsigned long long seed, unsigned long long idx) {
  const double u1 = unif(seed, idx, 0), u2 = unif(seed, idx, 1);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Source column of a data-domain column: a cyclic halo column is the wrap of
// the compute domain; false for halo columns/rows of a closed boundary.
__device__ inline bool src_col(const KDom &d, int ii, int jj, int &si, int &sj) {
  si = ii;
  sj = jj;
  if (jj < 1 || jj > d.nj) {
    if (!d.cyclic_y) return false;
    sj = 1 + wrapi(jj - 1, d.nj);
  }
  if (ii < 1 || ii > d.ni) {
    if (!d.cyclic_x) return false;
    si = 1 + wrapi(ii - 1, d.ni);
  }
  return true;
}

template <typename TT>
__global__ void k_synth_tracer(KDom d, TT *__restrict__ out, int kind, unsigned long long seed,
                               const double *__restrict__ zt, const int *__restrict__ kmt) {
  const int ii = blockIdx.x * SB + threadIdx.x, jj = blockIdx.y, k = blockIdx.z;
  if (ii >= d.NI) return;
  const size_t o = (size_t)ii + (size_t)d.NI * jj + (size_t)d.P2 * k;
  int si, sj;
  double v = 0.0;
  if (src_col(d, ii, jj, si, sj)) {
    const int os = si + d.NI * sj;
    if (k + 1 <= kmt[os]) {
      const unsigned long long idx = (unsigned long long)os + (unsigned long long)d.P2 * k;
      const double n = normal(seed, idx);
      if (kind == MOM_SYNTH_TEMP) v = 2.0 + 25.0 * exp(-zt[k] / 800.0) + 0.5 * n;
      else if (kind == MOM_SYNTH_SALT) v = 35.0 + 0.2 * n;
      else v = exp(log(1.0e-3) + 1.5 * n);
    }
  }
  out[o] = (TT)v;
}

__global__ void k_synth_logu(KDom d, double *__restrict__ out, int nplanes,
                             unsigned long long seed, double llo, double lhi,
                             const int *__restrict__ kmt) {
  const int ii = blockIdx.x * SB + threadIdx.x, jj = blockIdx.y, k = blockIdx.z;
  if (ii >= d.NI) return;
  int si, sj;
  const bool ok = src_col(d, ii, jj, si, sj);
  const int os = ok ? si + d.NI * sj : 0;
  for (int p = 0; p < nplanes; ++p) {
    const size_t lev = (size_t)k + (size_t)d.nk * p;
    double v = 0.0;
    if (ok && k + 1 <= kmt[os]) {
      const unsigned long long idx = (unsigned long long)os + (unsigned long long)d.P2 * lev;
      v = pow(10.0, llo + (lhi - llo) * unif(seed, idx, 2));
    }
    out[(size_t)ii + (size_t)d.NI * jj + (size_t)d.P2 * lev] = v;
  }
}

__global__ void k_synth_profile(KDom d, double *__restrict__ out, const double *__restrict__ prof) {
  const int ii = blockIdx.x * SB + threadIdx.x, jj = blockIdx.y, k = blockIdx.z;
  if (ii >= d.NI) return;
  out[(size_t)ii + (size_t)d.NI * jj + (size_t)d.P2 * k] = prof[k];
}

// Unit-amplitude face transports of level k from the streamfunction psi
// (U points) and the overturning amplitude ov (x faces):
//   du = coef_k (psik(j) - psik(j-1)),  dv = -coef_k (psik(i) - psik(i-1)),
//   psik = psi decay_k [k < kmu],  mov = rho0 (phi_k - phi_k+1),
//   phi_m = [m < kface] ov sin(pi zint_m / hden).
// prof = coef(nk), decay(nk), zint(nk+1).
__global__ void k_syn_uv(KDom d, const double *__restrict__ psi, const double *__restrict__ ov,
                         const double *__restrict__ hden, const int *__restrict__ kface,
                         const int *__restrict__ kmu, const double *__restrict__ prof,
                         double rho0, double *__restrict__ du, double *__restrict__ dv,
                         double *__restrict__ mov) {
  const int ii = blockIdx.x * SB + threadIdx.x, jj = blockIdx.y, k = blockIdx.z;
  if (ii >= d.NI) return;
  const int nk = d.nk;
  const double coef = prof[k], dec = prof[nk + k];
  const double z0 = prof[2 * nk + k], z1 = prof[2 * nk + k + 1];
  const int o = ii + d.NI * jj;
  const size_t ok = (size_t)o + (size_t)d.P2 * k;
  const double pc = psi[o] * dec * (kmu[o] >= k + 1 ? 1.0 : 0.0);
  double x = 0.0, y = 0.0;
  if (jj >= 1) {
    const int os = o - d.NI;
    x = coef * (pc - psi[os] * dec * (kmu[os] >= k + 1 ? 1.0 : 0.0));
  }
  if (ii >= 1) {
    const int ow = o - 1;
    y = -coef * (pc - psi[ow] * dec * (kmu[ow] >= k + 1 ? 1.0 : 0.0));
  }
  du[ok] = x;
  dv[ok] = y;
  const double pi = 3.141592653589793;
  const int kf = kface[o];
  const double ph0 = k < kf ? ov[o] * sin(pi * z0 / hden[o]) : 0.0;
  const double ph1 = k + 1 < kf ? ov[o] * sin(pi * z1 / hden[o]) : 0.0;
  mov[ok] = rho0 * (ph0 - ph1);
}

__device__ inline double tmk(const int *kmt, int o, int k) { return kmt[o] >= k + 1 ? 1.0 : 0.0; }

// du += mov*fac, then both transports masked to faces between wet cells;
// du zero on the first and last rows, dv on the last row.
__global__ void k_syn_combine(KDom d, double *__restrict__ du, double *__restrict__ dv,
                              const double *__restrict__ mov, double fac,
                              const int *__restrict__ kmt) {
  const int ii = blockIdx.x * SB + threadIdx.x, jj = blockIdx.y, k = blockIdx.z;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  const size_t ok = (size_t)o + (size_t)d.P2 * k;
  const double t = tmk(kmt, o, k);
  const double te = ii + 1 < d.NI ? tmk(kmt, o + 1, k) : 0.0;
  const double tn = jj + 1 < d.NJ ? tmk(kmt, o + d.NI, k) : 0.0;
  double x = du[ok] + mov[ok] * fac;
  x = x * (t * te);
  if (jj == 0 || jj == d.NJ - 1) x = 0.0;
  double y = dv[ok] * (t * tn);
  if (jj == d.NJ - 1) y = 0.0;
  du[ok] = x;
  dv[ok] = y;
}

// halo columns: wrap of the compute domain (cyclic) or zero (closed)
__global__ void k_syn_wrap(KDom d, double *__restrict__ a, int nplanes) {
  const int jj = blockIdx.x * SB + threadIdx.x, p = blockIdx.y;
  if (jj >= d.NJ || p >= nplanes) return;
  double *row = a + (size_t)d.P2 * p + (size_t)d.NI * jj;
  if (d.cyclic_x) {
    row[0] = row[d.NI - 2];
    row[d.NI - 1] = row[1];
  } else {
    row[0] = 0.0;
    row[d.NI - 1] = 0.0;
  }
}

// Largest advective Courant number of one column for transports scale*du,
// scale*dv and the w that continuity gives them.
__global__ void k_syn_cfl(KDom d, const double *__restrict__ du, const double *__restrict__ dv,
                          const double *__restrict__ dyte, const double *__restrict__ dxtn,
                          const double *__restrict__ dxte, const double *__restrict__ dytn,
                          const double *__restrict__ datr, const int *__restrict__ kmt,
                          const double *__restrict__ rdz, double scale, double dtt,
                          double *__restrict__ colmax) {
  const int ii = blockIdx.x * SB + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  double c = 0.0, w = 0.0;
  for (int k = 0; k < d.nk; ++k) {
    const size_t ok = (size_t)o + (size_t)d.P2 * k;
    const double rd = rdz[k];
    const double u = scale * du[ok] / dyte[o];
    if (ii < d.NI - 1) c = fmax(c, fabs(u * dtt * 2.0 / ((rd + rd) * dxte[o])));
    const double v = scale * dv[ok] / dxtn[o];
    if (jj < d.NJ - 1) c = fmax(c, fabs(v * dtt * 2.0 / ((rd + rd) * dytn[o])));
    const double t = tmk(kmt, o, k);
    double div = 0.0;
    if (ii >= 1) div = div + (scale * du[ok] - scale * du[ok - 1]);
    if (jj >= 1) div = div + (scale * dv[ok] - scale * dv[ok - d.NI]);
    div = div * (datr[o] * t);
    w = (div + w) * t;
    if (jj != 0 && jj != d.NJ - 1) c = fmax(c, fabs(w * dtt / rd));
  }
  colmax[o] = c;
}

// a(i,j,k) = scale * a(i,j,k) / met(i,j)
__global__ void k_syn_scale(KDom d, double *__restrict__ a, const double *__restrict__ met,
                            double scale) {
  const int ii = blockIdx.x * SB + threadIdx.x, jj = blockIdx.y, k = blockIdx.z;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  const size_t ok = (size_t)o + (size_t)d.P2 * k;
  a[ok] = scale * a[ok] / met[o];
}

// wrho_bt(0:nk) by continuity from uhrho_et, vhrho_nt (per column)
__global__ void k_syn_w(KDom d, const double *__restrict__ uh, const double *__restrict__ vh,
                        const double *__restrict__ dyte, const double *__restrict__ dxtn,
                        const double *__restrict__ datr, const int *__restrict__ kmt,
                        double *__restrict__ w) {
  const int ii = blockIdx.x * SB + threadIdx.x, jj = blockIdx.y;
  if (ii >= d.NI) return;
  const int o = ii + d.NI * jj;
  const bool edge_row = jj == 0 || jj == d.NJ - 1;
  double wk = 0.0;
  w[o] = 0.0;
  for (int k = 0; k < d.nk; ++k) {
    const size_t ok = (size_t)o + (size_t)d.P2 * k;
    double div = 0.0;
    if (ii >= 1) div = div + (uh[ok] * dyte[o] - uh[ok - 1] * dyte[o - 1]);
    if (jj >= 1) div = div + (vh[ok] * dxtn[o] - vh[ok - d.NI] * dxtn[o - d.NI]);
    const double t = tmk(kmt, o, k);
    div = div * (datr[o] * t);
    wk = (div + wk) * t;
    w[ok + d.P2] = edge_row ? 0.0 : wk;
  }
}

__global__ __launch_bounds__(SB) void k_syn_maxabs(const double *__restrict__ a, size_t n,
                                                   double *__restrict__ part) {
  __shared__ double sh[SB];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * SB + threadIdx.x; i < n; i += (size_t)gridDim.x * SB)
    m = fmax(m, fabs(a[i]));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = SB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

}  // namespace

#define SREQ(ctx, cond, ...)                                               \
  do {                                                                     \
    if (!(cond)) return mom_fail((ctx), MOM_GPU_ERR_ARG, __VA_ARGS__);     \
  } while (0)

namespace {

// a small host array copied to a temporary device buffer on ctx's stream
struct DevTmp {
  double *p = nullptr;
  ~DevTmp() {
    if (p) (void)hipFree(p);
  }
};

int upload_tmp(mom_gpu_ctx *ctx, DevTmp &t, const double *h, size_t n) {
  MOM_HIP(ctx, hipMalloc((void **)&t.p, sizeof(double) * n));
  MOM_HIP(ctx, hipMemcpyAsync(t.p, h, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  return MOM_GPU_OK;
}

dim3 grid3(const KDom &d, int nplanes) { return dim3((d.NI + SB - 1) / SB, d.NJ, nplanes); }

}  // namespace

extern "C" {

int mom_synth_tracer(mom_gpu_ctx_t *ctx, void *out, int dtype, int kind, unsigned long long seed,
                     const double *zt) {
  if (ctx) side_join(ctx);
  SREQ(nullptr, ctx, "mom_synth_tracer: NULL context");
  SREQ(ctx, out && zt, "mom_synth_tracer: NULL argument");
  SREQ(ctx, dtype == 8 || dtype == 4, "mom_synth_tracer: dtype must be 8 or 4");
  SREQ(ctx, kind >= 0 && kind <= 2, "mom_synth_tracer: unknown kind %d", kind);
  const KDom &d = ctx->kd;
  DevTmp z;
  int rc = upload_tmp(ctx, z, zt, d.nk);
  if (rc) return rc;
  if (dtype == 8)
    hipLaunchKernelGGL(k_synth_tracer<double>, grid3(d, d.nk), dim3(SB), 0, ctx->stream, d,
                       (double *)out, kind, seed, z.p, ctx->kmt);
  else
    hipLaunchKernelGGL(k_synth_tracer<float>, grid3(d, d.nk), dim3(SB), 0, ctx->stream, d,
                       (float *)out, kind, seed, z.p, ctx->kmt);
  MOM_LAUNCHED(ctx, "k_synth_tracer");
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MOM_GPU_OK;
}

int mom_synth_loguniform(mom_gpu_ctx_t *ctx, double *out, int nplanes, unsigned long long seed,
                         double lo, double hi) {
  if (ctx) side_join(ctx);
  SREQ(nullptr, ctx, "mom_synth_loguniform: NULL context");
  SREQ(ctx, out && nplanes >= 1 && lo > 0.0 && hi >= lo, "mom_synth_loguniform: bad argument");
  const KDom &d = ctx->kd;
  hipLaunchKernelGGL(k_synth_logu, grid3(d, d.nk), dim3(SB), 0, ctx->stream, d, out, nplanes,
                     seed, log10(lo), log10(hi), ctx->kmt);
  MOM_LAUNCHED(ctx, "k_synth_logu");
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MOM_GPU_OK;
}

int mom_synth_profile(mom_gpu_ctx_t *ctx, double *out, const double *prof, int nlev) {
  if (ctx) side_join(ctx);
  SREQ(nullptr, ctx, "mom_synth_profile: NULL context");
  SREQ(ctx, out && prof && nlev >= 1, "mom_synth_profile: bad argument");
  const KDom &d = ctx->kd;
  DevTmp p;
  int rc = upload_tmp(ctx, p, prof, nlev);
  if (rc) return rc;
  hipLaunchKernelGGL(k_synth_profile, grid3(d, nlev), dim3(SB), 0, ctx->stream, d, out, p.p);
  MOM_LAUNCHED(ctx, "k_synth_profile");
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MOM_GPU_OK;
}

int mom_synth_transports_unit(mom_gpu_ctx_t *ctx, const double *psi, const double *ov,
                              const double *hden, const int *kface, const double *prof,
                              double rho0, double *du, double *dv, double *mov) {
  if (ctx) side_join(ctx);
  SREQ(nullptr, ctx, "mom_synth_transports_unit: NULL context");
  SREQ(ctx, psi && ov && hden && kface && prof && du && dv && mov,
       "mom_synth_transports_unit: NULL argument");
  const KDom &d = ctx->kd;
  DevTmp p;
  int rc = upload_tmp(ctx, p, prof, 3 * (size_t)d.nk + 1);
  if (rc) return rc;
  hipLaunchKernelGGL(k_syn_uv, grid3(d, d.nk), dim3(SB), 0, ctx->stream, d, psi, ov, hden, kface,
                     ctx->kmu, p.p, rho0, du, dv, mov);
  MOM_LAUNCHED(ctx, "k_syn_uv");
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MOM_GPU_OK;
}

int mom_synth_maxabs(mom_gpu_ctx_t *ctx, const double *a, size_t n, double *out) {
  if (ctx) side_join(ctx);
  SREQ(nullptr, ctx, "mom_synth_maxabs: NULL context");
  SREQ(ctx, a && out && n > 0, "mom_synth_maxabs: bad argument");
  SREQ(ctx, ctx->red_n <= 1024, "mom_synth_maxabs: reduction buffer larger than expected");
  const int nb = ctx->red_n;  // partials fit the context's reduction buffer
  hipLaunchKernelGGL(k_syn_maxabs, dim3(nb), dim3(SB), 0, ctx->stream, a, n, ctx->red);
  MOM_LAUNCHED(ctx, "k_syn_maxabs");
  double part[1024];
  MOM_HIP(ctx, hipMemcpyAsync(part, ctx->red, sizeof(double) * nb, hipMemcpyDeviceToHost,
                              ctx->stream));
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double m = 0.0;
  for (int b = 0; b < nb; ++b) m = part[b] > m ? part[b] : m;
  *out = m;
  return MOM_GPU_OK;
}

int mom_synth_transports_combine(mom_gpu_ctx_t *ctx, double *du, double *dv, const double *mov,
                                 double fac) {
  if (ctx) side_join(ctx);
  SREQ(nullptr, ctx, "mom_synth_transports_combine: NULL context");
  SREQ(ctx, du && dv && mov, "mom_synth_transports_combine: NULL argument");
  const KDom &d = ctx->kd;
  hipLaunchKernelGGL(k_syn_combine, grid3(d, d.nk), dim3(SB), 0, ctx->stream, d, du, dv, mov, fac,
                     ctx->kmt);
  MOM_LAUNCHED(ctx, "k_syn_combine");
  const dim3 gw((d.NJ + SB - 1) / SB, d.nk);
  hipLaunchKernelGGL(k_syn_wrap, gw, dim3(SB), 0, ctx->stream, d, du, d.nk);
  MOM_LAUNCHED(ctx, "k_syn_wrap");
  hipLaunchKernelGGL(k_syn_wrap, gw, dim3(SB), 0, ctx->stream, d, dv, d.nk);
  MOM_LAUNCHED(ctx, "k_syn_wrap");
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MOM_GPU_OK;
}

int mom_synth_cfl(mom_gpu_ctx_t *ctx, const double *du, const double *dv, const double *rdz,
                  double scale, double dtt, double *colmax, double *out) {
  if (ctx) side_join(ctx);
  SREQ(nullptr, ctx, "mom_synth_cfl: NULL context");
  SREQ(ctx, du && dv && rdz && colmax && out, "mom_synth_cfl: NULL argument");
  const KDom &d = ctx->kd;
  DevTmp r;
  int rc = upload_tmp(ctx, r, rdz, d.nk);
  if (rc) return rc;
  hipLaunchKernelGGL(k_syn_cfl, dim3((d.NI + SB - 1) / SB, d.NJ), dim3(SB), 0, ctx->stream, d, du,
                     dv, ctx->dyte, ctx->dxtn, ctx->dxte, ctx->dytn, ctx->datr, ctx->kmt, r.p,
                     scale, dtt, colmax);
  MOM_LAUNCHED(ctx, "k_syn_cfl");
  return mom_synth_maxabs(ctx, colmax, (size_t)d.P2, out);
}

int mom_synth_transports_finish(mom_gpu_ctx_t *ctx, double *du, double *dv, double scale,
                                double *wrho) {
  if (ctx) side_join(ctx);
  SREQ(nullptr, ctx, "mom_synth_transports_finish: NULL context");
  SREQ(ctx, du && dv && wrho, "mom_synth_transports_finish: NULL argument");
  const KDom &d = ctx->kd;
  hipLaunchKernelGGL(k_syn_scale, grid3(d, d.nk), dim3(SB), 0, ctx->stream, d, du, ctx->dyte,
                     scale);
  MOM_LAUNCHED(ctx, "k_syn_scale");
  hipLaunchKernelGGL(k_syn_scale, grid3(d, d.nk), dim3(SB), 0, ctx->stream, d, dv, ctx->dxtn,
                     scale);
  MOM_LAUNCHED(ctx, "k_syn_scale");
  hipLaunchKernelGGL(k_syn_w, dim3((d.NI + SB - 1) / SB, d.NJ), dim3(SB), 0, ctx->stream, d, du,
                     dv, ctx->dyte, ctx->dxtn, ctx->datr, ctx->kmt, wrho);
  MOM_LAUNCHED(ctx, "k_syn_w");
  hipLaunchKernelGGL(k_syn_wrap, dim3((d.NJ + SB - 1) / SB, d.nk + 1), dim3(SB), 0, ctx->stream,
                     d, wrho, d.nk + 1);
  MOM_LAUNCHED(ctx, "k_syn_wrap");
  MOM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MOM_GPU_OK;
}

}  // extern "C"
