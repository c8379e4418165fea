// pxr_mem.hip -- device memory, copies, fills, timers and the energy
// reductions of the diagnostics, all on the library's own HIP runtime and
// stream.  Callers that hold no other GPU runtime (C hosts, the Python
// manager layer) allocate and move data through these entry points.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "pxr_device.h"
#include "pxr_runtime.h"

using namespace pxr;

namespace pxr {

__global__ __launch_bounds__(256) void k_fill_f64(double *p, i64 n, double v) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_i64(int64_t *p, i64 n, int64_t v) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// plain device copy in 16-byte words (grid-stride); also the reference
// point for HBM traffic counters: it reads and writes each byte once
__global__ __launch_bounds__(256) void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src,
                                                i64 n) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// deterministic two-level sums: fixed blocks of RED_B elements, then one
// workgroup over the block partials
#define RED_B 4096
__device__ __forceinline__ double block_sum(double v, double *s) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  return s[0];
}

// sum over particles of (1/gaminv - 1) * w  (kin_energy.F90:33-143,
// diags.F90:908-911; multiplied by m c^2 on the host)
__global__ __launch_bounds__(256) void k_kin_partial(i64 np, const double *gi, const double *w,
                                                     double *part) {
  __shared__ double s[256];
  const i64 b0 = (i64)blockIdx.x * RED_B;
  double acc = 0.0;
  for (int k = threadIdx.x; k < RED_B; k += 256) {
    const i64 i = b0 + k;
    if (i < np) acc += (1.0 / gi[i] - 1.0) * w[i];
  }
  const double t = block_sum(acc, s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// gamma - 1 = u^2 / (1 + sqrt(1 + u^2)) in FP64 from FP32 momenta: an FP32
// gaminv is too coarse for the kinetic energy of a cold species (for
// u ~ 1e-3, 1/gaminv - 1 ~ 5e-7 is a few FP32 ulps, and the rounding of an
// FP32 pusher's 1 + u^2 biases the sum by percents)
__global__ __launch_bounds__(256) void k_kin_partial_f32(i64 np, const float *ux, const float *uy,
                                                         const float *uz, const float *w,
                                                         double *part) {
  __shared__ double s[256];
  const i64 b0 = (i64)blockIdx.x * RED_B;
  double acc = 0.0;
  for (int k = threadIdx.x; k < RED_B; k += 256) {
    const i64 i = b0 + k;
    if (i < np) {
      const double a = ux[i], b = uy[i], c = uz[i];
      const double usq = a * a + b * b + c * c;
      acc += usq / (1.0 + sqrt(1.0 + usq)) * (double)w[i];
    }
  }
  const double t = block_sum(acc, s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// sum of f^2 over the nodes 0..n-1 of each axis (diags.F90:1072-1114)
__global__ __launch_bounds__(256) void k_sq_partial(const double *f, GridGeom g, double *part) {
  __shared__ double s[256];
  const i64 tot = (i64)g.nx * g.ny * g.nz;
  const i64 b0 = (i64)blockIdx.x * RED_B;
  double acc = 0.0;
  for (int k = threadIdx.x; k < RED_B; k += 256) {
    const i64 c = b0 + k;
    if (c < tot) {
      const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), l = (int)(c / ((i64)g.nx * g.ny));
      const double v = f[g.idx(i, j, l)];
      acc += v * v;
    }
  }
  const double t = block_sum(acc, s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_sum_final(const double *part, i64 n, double *out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (i64 i = threadIdx.x; i < n; i += 256) acc += part[i];
  const double t = block_sum(acc, s);
  if (threadIdx.x == 0) *out = t;
}

static std::vector<hipEvent_t> g_events;

static int reduce_finish(double *part, i64 nblk, double *out) {
  double *res = part + nblk;
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, rt().stream, part, nblk, res);
  PXR_LAUNCHED();
  PXR_HIP(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, rt().stream));
  PXR_HIP(hipStreamSynchronize(rt().stream));
  return PXRHIP_OK;
}

}  // namespace pxr

extern "C" {

int pxrhip_malloc(void **ptr, int64_t bytes) {
  PXR_INIT();
  if (!ptr || bytes < 0) return set_error(PXRHIP_ERR_ARG, "bad malloc request");
  *ptr = nullptr;
  if (bytes == 0) bytes = 8;
  PXR_HIP(hipMalloc(ptr, (size_t)bytes));
  return PXRHIP_OK;
}

int pxrhip_free(void *ptr) {
  PXR_INIT();
  if (!ptr) return PXRHIP_OK;
  PXR_HIP(hipStreamSynchronize(rt().stream));
  PXR_HIP(hipFree(ptr));
  return PXRHIP_OK;
}

int pxrhip_memcpy_htod(void *dst, const void *src, int64_t bytes) {
  PXR_INIT();
  if (bytes < 0 || (bytes && (!dst || !src))) return set_error(PXRHIP_ERR_ARG, "bad copy");
  if (!bytes) return PXRHIP_OK;
  PXR_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, rt().stream));
  PXR_HIP(hipStreamSynchronize(rt().stream));  // the host buffer may be reused on return
  return PXRHIP_OK;
}

int pxrhip_memcpy_dtoh(void *dst, const void *src, int64_t bytes) {
  PXR_INIT();
  if (bytes < 0 || (bytes && (!dst || !src))) return set_error(PXRHIP_ERR_ARG, "bad copy");
  if (!bytes) return PXRHIP_OK;
  PXR_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, rt().stream));
  PXR_HIP(hipStreamSynchronize(rt().stream));
  return PXRHIP_OK;
}

int pxrhip_memcpy_dtod(void *dst, const void *src, int64_t bytes) {
  PXR_INIT();
  if (bytes < 0 || (bytes && (!dst || !src))) return set_error(PXRHIP_ERR_ARG, "bad copy");
  if (!bytes) return PXRHIP_OK;
  if ((((uintptr_t)dst | (uintptr_t)src | (uintptr_t)bytes) & 15) == 0) {
    const i64 n = bytes / 16;
    const i64 nblk = std::min<i64>((n + 255) / 256, (i64)rt().num_cus * 16);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)nblk), dim3(256), 0, rt().stream, (uint4 *)dst,
                       (const uint4 *)src, n);
    PXR_LAUNCHED();
    return PXRHIP_OK;
  }
  PXR_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, rt().stream));
  return PXRHIP_OK;
}

int pxrhip_fill_f64(double *p, int64_t n, double value) {
  PXR_INIT();
  if (n < 0 || (n && !p)) return set_error(PXRHIP_ERR_ARG, "bad fill");
  if (!n) return PXRHIP_OK;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid1d(n, 256)), dim3(256), 0, rt().stream, p, n, value);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_fill_i64(int64_t *p, int64_t n, int64_t value) {
  PXR_INIT();
  if (n < 0 || (n && !p)) return set_error(PXRHIP_ERR_ARG, "bad fill");
  if (!n) return PXRHIP_OK;
  hipLaunchKernelGGL(k_fill_i64, dim3(grid1d(n, 256)), dim3(256), 0, rt().stream, p, n, value);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_timer_record(int slot) {
  PXR_INIT();
  if (slot < 0 || slot > 65535) return set_error(PXRHIP_ERR_ARG, "timer slot must be 0..65535");
  while ((int)g_events.size() <= slot) {
    hipEvent_t e;
    PXR_HIP(hipEventCreate(&e));
    g_events.push_back(e);
  }
  PXR_HIP(hipEventRecord(g_events[slot], rt().stream));
  return PXRHIP_OK;
}

int pxrhip_timer_elapsed(int start, int stop, double *ms) {
  PXR_INIT();
  if (start < 0 || stop < 0 || start >= (int)g_events.size() || stop >= (int)g_events.size() || !ms)
    return set_error(PXRHIP_ERR_ARG, "timer slot never recorded");
  PXR_HIP(hipEventSynchronize(g_events[stop]));
  float f = 0.f;
  PXR_HIP(hipEventElapsedTime(&f, g_events[start], g_events[stop]));
  *ms = f;
  return PXRHIP_OK;
}

int pxrhip_kinetic_energy(int64_t np, const double *gaminv, const double *w, double *out) {
  PXR_INIT();
  if (np < 0 || !out || (np && (!gaminv || !w))) return set_error(PXRHIP_ERR_ARG, "bad arguments");
  *out = 0.0;
  if (!np) return PXRHIP_OK;
  const i64 nblk = (np + RED_B - 1) / RED_B;
  double *part = (double *)scratch(11, sizeof(double) * (size_t)(nblk + 1));
  if (!part) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
  hipLaunchKernelGGL(k_kin_partial, dim3((unsigned)nblk), dim3(256), 0, rt().stream, np, gaminv, w,
                     part);
  PXR_LAUNCHED();
  return reduce_finish(part, nblk, out);
}

int pxrhip_kinetic_energy_f32(int64_t np, const float *ux, const float *uy, const float *uz,
                              const float *w, double *out) {
  PXR_INIT();
  if (np < 0 || !out || (np && (!ux || !uy || !uz || !w)))
    return set_error(PXRHIP_ERR_ARG, "bad arguments");
  *out = 0.0;
  if (!np) return PXRHIP_OK;
  const i64 nblk = (np + RED_B - 1) / RED_B;
  double *part = (double *)scratch(11, sizeof(double) * (size_t)(nblk + 1));
  if (!part) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
  hipLaunchKernelGGL(k_kin_partial_f32, dim3((unsigned)nblk), dim3(256), 0, rt().stream, np, ux,
                     uy, uz, w, part);
  PXR_LAUNCHED();
  return reduce_finish(part, nblk, out);
}

int pxrhip_field_energy(const double *f, int64_t nx, int64_t ny, int64_t nz, int64_t ngx,
                        int64_t ngy, int64_t ngz, double *out) {
  PXR_INIT();
  if (!f || !out || nx < 1 || ny < 1 || nz < 1 || ngx < 0 || ngy < 0 || ngz < 0)
    return set_error(PXRHIP_ERR_ARG, "bad arguments");
  GridGeom g;
  g.nx = (int)nx; g.ny = (int)ny; g.nz = (int)nz;
  g.gx = (int)ngx; g.gy = (int)ngy; g.gz = (int)ngz;
  g.NX = nx + 1 + 2 * ngx;
  g.NY = ny + 1 + 2 * ngy;
  const i64 tot = nx * ny * nz;
  const i64 nblk = (tot + RED_B - 1) / RED_B;
  double *part = (double *)scratch(11, sizeof(double) * (size_t)(nblk + 1));
  if (!part) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
  hipLaunchKernelGGL(k_sq_partial, dim3((unsigned)nblk), dim3(256), 0, rt().stream, f, g, part);
  PXR_LAUNCHED();
  return reduce_finish(part, nblk, out);
}

}  // extern "C"
