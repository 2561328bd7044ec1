/*
 * aco_noalloc.hip - route costs for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).
 *
 * aco_hip_route_costs stages the caller's tours, lengths and costs in three
 * device buffers it allocates per call.  aco_hip_route_costs_noalloc gives
 * the same result through `scratch`, a fixed device array of
 * ACO_HIP_NOALLOC_SCRATCH_BYTES: every GPU the library runs on holds its own
 * copy until the process ends, and every call on that GPU reuses it, so a
 * call allocates nothing.  The kernel writes every cost it reports, so
 * nothing in the array needs clearing between calls; calls that overlap on
 * one GPU (from several host threads) race on the array.
 *
 * `scratch` names a host-side stand-in, not a device address: host code
 * uses only the address hipGetSymbolAddress returned for the device, looked
 * up once per device while that device is current (noalloc_attach_device,
 * called by aco_hip_create).
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "aco_ctx.h"

static __device__ __align__(256) unsigned char scratch[ACO_HIP_NOALLOC_SCRATCH_BYTES];

#define NOALLOC_MAX_DEVICES 64

static std::mutex g_mu;
static unsigned char *g_ptr[NOALLOC_MAX_DEVICES];   /* device address of scratch */
static hipError_t g_err[NOALLOC_MAX_DEVICES];        /* why the lookup failed */

int noalloc_attach_device(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return -1;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[device]) return 0;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
    return -1;
  }
  g_ptr[device] = static_cast<unsigned char *>(p);
  return 0;
}

static size_t r256(size_t x) { return (x + 255) & ~(size_t)255; }

extern "C" int aco_hip_route_costs_noalloc(AcoHipContext *c, int inst, const int32_t *tours,
                                           const int32_t *len, int count, int stride,
                                           int64_t *cost) {
  NEED_LOADED(c);
  int rc = check_route_cost_args(c, inst, tours, len, count, stride, cost);
  if (rc || count <= 0) return rc;
  const size_t tb = r256(sizeof(int32_t) * (size_t)count * (size_t)stride);
  const size_t lb = r256(sizeof(int32_t) * (size_t)count);
  const size_t cb = sizeof(long long) * (size_t)count;
  if (tb + lb + cb > (size_t)ACO_HIP_NOALLOC_SCRATCH_BYTES)
    return set_err(c, ACO_HIP_ERR_INVALID_INPUT,
                   "route-cost batch needs %zu bytes of staging, more than the %u bytes of the "
                   "allocation-free entry point",
                   tb + lb + cb, (unsigned)ACO_HIP_NOALLOC_SCRATCH_BYTES);
  unsigned char *base = nullptr;
  hipError_t why = hipSuccess;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) {
      base = g_ptr[c->device];
      why = g_err[c->device];
    }
  }
  if (!base)
    return set_err(c, ACO_HIP_ERR_BACKEND, "no device address for the staging array: %s",
                   hipGetErrorString(why));
  int32_t *dt = reinterpret_cast<int32_t *>(base);
  int32_t *dl = reinterpret_cast<int32_t *>(base + tb);
  long long *dc = reinterpret_cast<long long *>(base + tb + lb);
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipMemcpyAsync(dt, tours, sizeof(int32_t) * (size_t)count * stride,
                         hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(dl, len, sizeof(int32_t) * count, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_route_costs, dim3((count * 64 + 255) / 256), dim3(256), 0, c->stream, c->P,
                     inst, dt, dl, count, stride, dc);
  LAUNCHED(c, "k_route_costs");
  HIPC(c, hipMemcpyAsync(cost, dc, cb, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return ACO_HIP_OK;
}
