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

__global__ __launch_bounds__(256) void k_f32_to_f64(const float *s, double *d, i64 n) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (double)s[i];
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

// divergence diagnostics, one thread per node (x fastest):
// calc_field_div (diags.F90:129-156), backward differences of E at the nodes
// 0..n of each axis; calc_field_divB (diags.F90:170-197), forward
// differences of B at the nodes 1..n.  The sum is grouped as the reference's
// (x + y) + z.
template <bool FORWARD>
__global__ __launch_bounds__(256) void k_div(double *__restrict__ out, const double *__restrict__ fx,
                                             const double *__restrict__ fy,
                                             const double *__restrict__ fz, GridGeom g, int lo,
                                             int mx, int my, i64 tot, double idx, double idy,
                                             double idz) {
  const i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= tot) return;
  const int i = lo + (int)(c % mx), j = lo + (int)((c / mx) % my), l = lo + (int)(c / ((i64)mx * my));
  const i64 p = g.idx(i, j, l);
  const i64 sy = g.NX, sz = g.NX * g.NY;
  if (FORWARD)
    out[p] = idx * (fx[p + 1] - fx[p]) + idy * (fy[p + sy] - fy[p]) + idz * (fz[p + sz] - fz[p]);
  else
    out[p] = idx * (fx[p] - fx[p - 1]) + idy * (fy[p] - fy[p - sy]) + idz * (fz[p] - fz[p - sz]);
}

// sum over the nodes 1..n of each axis of (a*eps0 - b)^2, or of a^2 when b
// is null (get_loc_norm_divErho and get_loc_norm_2, diags.F90:1222-1293)
__global__ __launch_bounds__(256) void k_norm_partial(const double *a, const double *b,
                                                      double eps0, GridGeom g, double *part) {
  __shared__ double s[256];
  const i64 tot = (i64)g.nx * g.ny * g.nz;
  const i64 b0 = (i64)blockIdx.x * RED_B;
  double acc = 0.0;
  for (int k = threadIdx.x; k < RED_B; k += 256) {
    const i64 c = b0 + k;
    if (c < tot) {
      const int i = 1 + (int)(c % g.nx), j = 1 + (int)((c / g.nx) % g.ny),
                l = 1 + (int)(c / ((i64)g.nx * g.ny));
      const i64 p = g.idx(i, j, l);
      const double v = b ? a[p] * eps0 - b[p] : a[p];
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

int pxrhip_convert_f32_f64(const float *src, double *dst, int64_t n) {
  PXR_INIT();
  if (n < 0 || (n && (!src || !dst))) return set_error(PXRHIP_ERR_ARG, "bad conversion");
  if (!n) return PXRHIP_OK;
  hipLaunchKernelGGL(k_f32_to_f64, dim3(grid1d(n, 256)), dim3(256), 0, rt().stream, src, dst, n);
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

static bool grid_args(int64_t nx, int64_t ny, int64_t nz, int64_t ngx, int64_t ngy, int64_t ngz,
                      int64_t min_guard, GridGeom &g) {
  if (nx < 1 || ny < 1 || nz < 1 || ngx < min_guard || ngy < min_guard || ngz < min_guard)
    return false;
  g.nx = (int)nx; g.ny = (int)ny; g.nz = (int)nz;
  g.gx = (int)ngx; g.gy = (int)ngy; g.gz = (int)ngz;
  g.NX = nx + 1 + 2 * ngx;
  g.NY = ny + 1 + 2 * ngy;
  return true;
}

static int field_div(bool forward, double *out, const double *fx, const double *fy,
                     const double *fz, int64_t nx, int64_t ny, int64_t nz, int64_t ngx,
                     int64_t ngy, int64_t ngz, double dx, double dy, double dz) {
  PXR_INIT();
  GridGeom g;
  if (!out || !fx || !fy || !fz || !grid_args(nx, ny, nz, ngx, ngy, ngz, 1, g))
    return set_error(PXRHIP_ERR_ARG, "bad arguments (the differences need guards >= 1)");
  if (!(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0))
    return set_error(PXRHIP_ERR_ARG, "cell sizes must be positive");
  // nodes 0..n (E) or 1..n (B) of each axis
  const int lo = forward ? 1 : 0;
  const int mx = (int)nx + 1 - lo, my = (int)ny + 1 - lo, mz = (int)nz + 1 - lo;
  const i64 tot = (i64)mx * my * mz;
  const double idx = 1.0 / dx, idy = 1.0 / dy, idz = 1.0 / dz;
  if (forward)
    hipLaunchKernelGGL(k_div<true>, dim3(grid1d(tot, 256)), dim3(256), 0, rt().stream, out, fx, fy,
                       fz, g, lo, mx, my, tot, idx, idy, idz);
  else
    hipLaunchKernelGGL(k_div<false>, dim3(grid1d(tot, 256)), dim3(256), 0, rt().stream, out, fx,
                       fy, fz, g, lo, mx, my, tot, idx, idy, idz);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_calc_field_div(double *divee, const double *eex, const double *eey, const double *eez,
                          int64_t nx, int64_t ny, int64_t nz, int64_t nxguard, int64_t nyguard,
                          int64_t nzguard, double dx, double dy, double dz) {
  return field_div(false, divee, eex, eey, eez, nx, ny, nz, nxguard, nyguard, nzguard, dx, dy, dz);
}

int pxrhip_calc_field_divb(double *divbb, const double *bbx, const double *bby, const double *bbz,
                           int64_t nx, int64_t ny, int64_t nz, int64_t nxguard, int64_t nyguard,
                           int64_t nzguard, double dx, double dy, double dz) {
  return field_div(true, divbb, bbx, bby, bbz, nx, ny, nz, nxguard, nyguard, nzguard, dx, dy, dz);
}

static int node_norm(const double *a, const double *b, int64_t nx, int64_t ny, int64_t nz,
                     int64_t ngx, int64_t ngy, int64_t ngz, double *out) {
  PXR_INIT();
  GridGeom g;
  if (!a || !out || !grid_args(nx, ny, nz, ngx, ngy, ngz, 0, g))
    return set_error(PXRHIP_ERR_ARG, "bad arguments");
  const i64 tot = nx * ny * nz;
  const i64 nblk = (tot + RED_B - 1) / RED_B;
  double *part = (double *)scratch(11, sizeof(double) * (size_t)(nblk + 1));
  if (!part) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
  // eps0 = 1: normalised units, the reference's LIBRARY build (modules.F90)
  hipLaunchKernelGGL(k_norm_partial, dim3((unsigned)nblk), dim3(256), 0, rt().stream, a, b, 1.0, g,
                     part);
  PXR_LAUNCHED();
  return reduce_finish(part, nblk, out);
}

int pxrhip_get_norm_diverho(const double *divee2, const double *rho2, int64_t nx2, int64_t ny2,
                            int64_t nz2, int64_t nxguard, int64_t nyguard, int64_t nzguard,
                            double *norm) {
  if (!rho2) return set_error(PXRHIP_ERR_ARG, "rho2 is null");
  const int rc = node_norm(divee2, rho2, nx2, ny2, nz2, nxguard, nyguard, nzguard, norm);
  if (rc == PXRHIP_OK) *norm = sqrt(*norm);
  return rc;
}

int pxrhip_get_loc_norm_2(const double *array, int64_t nx2, int64_t ny2, int64_t nz2,
                          int64_t nxguard, int64_t nyguard, int64_t nzguard, double *norm) {
  return node_norm(array, nullptr, nx2, ny2, nz2, nxguard, nyguard, nzguard, norm);
}

}  // extern "C"
