// Masked tracer content with nothing allocated: the content kernels of
// mom_gpu_tracer_content (ocean_diag/ocean_tracer_diag.F90:2648-2661) with
// their partial sums in a fixed array of this file instead of the context's
// reduction buffer, and the result left in the caller's device memory.  For
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops).
//
// `scratch` is one array per GPU, alive until the process ends.  Host code
// never passes its name to a kernel (in host code that name is a stand-in,
// not a device address): the first call on a device looks up that device's
// address with hipGetSymbolAddress, with the device current, and every call
// after it passes that address.  All calls on one GPU share the array and
// each overwrites every partial it reads, so nothing needs clearing; calls
// that overlap on one GPU (other streams, other contexts) race.
#include "mom_internal.h"

#include <mutex>

// partial (hi, lo) pairs of up to NOALLOC_BLOCKS reduction blocks: 16 KiB
constexpr int NOALLOC_BLOCKS = 1024;
static __device__ __align__(256) double scratch[2 * NOALLOC_BLOCKS];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on `device` (looked up once per device).
int scratch_on(mom_gpu_ctx *ctx, int device, double **out) {
  if (device < 0 || device >= MAX_DEVICES)
    return mom_fail(ctx, MOM_GPU_ERR_ARG, "mom_gpu_tracer_content_noalloc: device %d outside "
                    "[0, %d)", device, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    int cur = -1;
    MOM_HIP(ctx, hipGetDevice(&cur));
    if (cur != device) MOM_HIP(ctx, hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) MOM_HIP(ctx, hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return mom_fail(ctx, MOM_GPU_ERR_HIP, "mom_gpu_tracer_content_noalloc: "
                      "hipGetSymbolAddress(scratch) on device %d: %s", device,
                      hipGetErrorString(e));
    g_ptr[device] = (double *)p;
  }
  *out = g_ptr[device];
  return MOM_GPU_OK;
}

}  // namespace

extern "C" int mom_gpu_tracer_content_noalloc(mom_gpu_ctx_t *ctx, const void *T, int dtype,
                                              const double *rho_dzt, double *content_dev) {
  if (!ctx)
    return mom_fail(nullptr, MOM_GPU_ERR_ARG, "mom_gpu_tracer_content_noalloc: NULL context");
  side_join(ctx);
  if (!T || !rho_dzt || !content_dev)
    return mom_fail(ctx, MOM_GPU_ERR_ARG, "mom_gpu_tracer_content_noalloc: NULL argument");
  if (dtype != 8 && dtype != 4)
    return mom_fail(ctx, MOM_GPU_ERR_ARG, "mom_gpu_tracer_content_noalloc: dtype %d is not 8 "
                    "(FP64) or 4 (FP32)", dtype);
  if (ctx->red_n > NOALLOC_BLOCKS)
    return mom_fail(ctx, MOM_GPU_ERR_ARG, "mom_gpu_tracer_content_noalloc: the reduction "
                    "uses %d blocks, more than the %d whose partial sums the fixed buffer "
                    "holds", ctx->red_n, NOALLOC_BLOCKS);
  double *part = nullptr;
  const int rc = scratch_on(ctx, ctx->device, &part);
  if (rc != MOM_GPU_OK) return rc;
  return launch_content_part(ctx, T, dtype, rho_dzt, part, content_dev);
}
