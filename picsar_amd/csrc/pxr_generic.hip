// pxr_generic.hip -- the reference's generic-order leaf routines on gfx950,
// with one shape order per axis:
//   pxr_gete3d_n_energy_conserving, pxr_getb3d_n_energy_conserving
//       (field_gathering_on_3d.F90:1042-1365, 697-1026): orders 1..3;
//   pxr_depose_jxjyjz_esirkepov_n (esirkepov_3d.F90:3480-3836): orders 0..3;
//   pxr_depose_rho_n (charge_deposition_3d.F90:2081-2228): orders 0..3.
// The dispatchers (pxr_leaf.hip) call them where the reference's do: for
// unequal orders, order 0, field_gathe_algo 2 and current_depo_algo 1.
//
// One particle per lane.  The orders are run-time values: every loop runs
// over the largest support (4 weights, 6 nodes of an Esirkepov box) with the
// order's bound tested inside, so that the unrolled weights stay in
// registers.  Products and sums follow the reference's operation order with
// no contraction, so the gather and each particle's current contributions
// are the reference's bit for bit; the grid sums of the deposition are FP64
// atomics, whose order varies.
#include "pxr_device.h"
#include "pxr_generic.h"
#include "pxr_runtime.h"

namespace pxr {

// The generic gather's constants onesixth = 1./6. and twothird = 2./3. are
// default-real (single-precision) literals widened to FP64
// (field_gathering_on_3d.F90:1071); the deposition routines' are FP64.
#define GN_SIXTH ((double)(1.0f / 6.0f))
#define GN_TWOTHIRD ((double)(2.0f / 3.0f))
#define ONESIXTH_D (1.0 / 6.0)
#define TWOTHIRD_D (2.0 / 3.0)
#define ONETHIRD_D (1.0 / 3.0)

// Weights of order n (0..3) at offset f from the base node, in the generic
// routines' evaluation order (esirkepov_3d.F90:3588-3608).
PXR_DEV void w_order(int n, double f, double *w, double sixth, double twothird) {
  if (n == 0) {
    w[0] = 1.0;
  } else if (n == 1) {
    w[0] = 1.0 - f;
    w[1] = f;
  } else if (n == 2) {
    const double a = 0.5 - f, b = 0.5 + f;
    w[0] = 0.5 * (a * a);
    w[1] = 0.75 - f * f;
    w[2] = 0.5 * (b * b);
  } else {
    const double o = 1.0 - f, f2 = f * f, o2 = o * o;
    w[0] = sixth * o2 * o;
    w[1] = twothird - f2 * (1.0 - f / 2.0);
    w[2] = twothird - o2 * (1.0 - o / 2.0);
    w[3] = sixth * f2 * f;
  }
}

// Node index of an order-n shape: nint for even orders, floor for odd ones.
PXR_DEV int n_index(int n, double x) { return (n & 1) ? f_floor(x) : f_nint(x); }

// ---------------------------------------------------------------------------
// gather
// ---------------------------------------------------------------------------
// One axis of the generic gather: the node-centred shape of order n (first
// node j - n/2) and the staggered one of order m = n - 1 (lower order in v)
// or n at j0 = floor(x - s) or floor(x + 1/2 - s) by the parity rules of
// field_gathering_on_3d.F90:1133-1167, offset (x - s) - j0.
struct GAxis {
  double sn[4], ss[4];
  int fn, fs, n, m;
};

PXR_DEV void gather_axis(GAxis &ax, int n, int lower, double s, double x) {
  const bool even = (n & 1) == 0;
  const int j = even ? f_nint(x) : f_floor(x);
  const int j0 = (even == (lower != 0)) ? f_floor(x - s) : f_floor((x + 0.5) - s);
  ax.n = n;
  ax.m = lower ? n - 1 : n;
  w_order(n, x - (double)j, ax.sn, GN_SIXTH, GN_TWOTHIRD);
  w_order(ax.m, (x - s) - (double)j0, ax.ss, GN_SIXTH, GN_TWOTHIRD);
  ax.fn = j - n / 2;
  ax.fs = j0 - (lower ? (n - 1) / 2 : n / 2);
}

// acc + sum_c sum_b sum_a wa[a] wb[b] wc[c] F(i0+a, j0+b, k0+c) * sgn, one
// term at a time in the reference's loop order (z outer, x inner).
PXR_DEV double gn_sum(double acc, const double *F, i64 sy, i64 sz, int i0, int j0, int k0,
                      const double *wa, int na, const double *wb, int nb, const double *wc,
                      int nc, double sgn) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
    if (c > nc) continue;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      if (b > nb) continue;
      const double *row = F + (i0 + sy * (i64)(j0 + b) + sz * (i64)(k0 + c));
#pragma unroll
      for (int a = 0; a < 4; a++) {
        if (a > na) continue;
        acc = acc + wa[a] * wb[b] * wc[c] * row[a] * sgn;
      }
    }
  }
  return acc;
}

__global__ __launch_bounds__(256) void k_gather_n(GatherNArgs a) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= a.np) return;
  double p[3] = {(a.xp[ip] - a.xmin) * a.dxi, (a.yp[ip] - a.ymin) * a.dyi,
                 (a.zp[ip] - a.zmin) * a.dzi};
  double sgx = 1.0, sgy = 1.0;
  if (a.l4symtry) {  // field_gathering_on_3d.F90:1120-1131
    if (p[0] < 0.0) { p[0] = -p[0]; sgx = -1.0; }
    if (p[1] < 0.0) { p[1] = -p[1]; sgy = -1.0; }
  }
  GAxis X, Y, Z;
  gather_axis(X, a.no[0], a.lower, a.s, p[0]);
  gather_axis(Y, a.no[1], a.lower, a.s, p[1]);
  gather_axis(Z, a.no[2], a.lower, a.s, p[2]);
  const int lo[3] = {min(X.fn, X.fs), min(Y.fn, Y.fs), min(Z.fn, Z.fs)};
  const int hi[3] = {max(X.fn + X.n, X.fs + X.m), max(Y.fn + Y.n, Y.fs + Y.m),
                     max(Z.fn + Z.n, Z.fs + Z.m)};
  if (!a.g.inside(lo[0], lo[1], lo[2]) || !a.g.inside(hi[0], hi[1], hi[2])) {
    count_bad(a.bad);
    return;
  }
  const i64 o = a.g.idx(0, 0, 0), sy = a.g.NX, sz = a.g.NX * a.g.NY;
  // E (pxr_gete3d_n_energy_conserving, :1336-1361), then B (:995-1021)
  a.ex[ip] = gn_sum(a.ex[ip], a.exg + o, sy, sz, X.fs, Y.fn, Z.fn, X.ss, X.m, Y.sn, Y.n, Z.sn,
                    Z.n, sgx);
  a.ey[ip] = gn_sum(a.ey[ip], a.eyg + o, sy, sz, X.fn, Y.fs, Z.fn, X.sn, X.n, Y.ss, Y.m, Z.sn,
                    Z.n, sgy);
  a.ez[ip] = gn_sum(a.ez[ip], a.ezg + o, sy, sz, X.fn, Y.fn, Z.fs, X.sn, X.n, Y.sn, Y.n, Z.ss,
                    Z.m, 1.0);
  a.bx[ip] = gn_sum(a.bx[ip], a.bxg + o, sy, sz, X.fn, Y.fs, Z.fs, X.sn, X.n, Y.ss, Y.m, Z.ss,
                    Z.m, sgx);
  a.by[ip] = gn_sum(a.by[ip], a.byg + o, sy, sz, X.fs, Y.fn, Z.fs, X.ss, X.m, Y.sn, Y.n, Z.ss,
                    Z.m, sgy);
  a.bz[ip] = gn_sum(a.bz[ip], a.bzg + o, sy, sz, X.fs, Y.fs, Z.fn, X.ss, X.m, Y.ss, Y.m, Z.sn,
                    Z.n, 1.0);
}

// The reference multiplies Ez and Bz by no sign; times 1.0 is exact, so the
// shared sum above changes no bit for them.

// ---------------------------------------------------------------------------
// Esirkepov current, pxr_depose_jxjyjz_esirkepov_n
// ---------------------------------------------------------------------------
// One axis: new-position weights s0 and differences ds = s(old) - s0(new)
// over the box nodes -2..3 around the new node (index r + 2), and the box
// [lo, hi] = [min(0,d) - n/2, max(0,d) + (n+1)/2].
struct EnAxis {
  double s0[6], ds[6];
  int base, lo, hi;
};

// weight r - first of ww[0..n] (0 outside), selected without indexing
PXR_DEV double pick(const double *ww, int n, int k) {
  double v = 0.0;
  v = (k == 0) ? ww[0] : v;
  v = (k == 1 && n >= 1) ? ww[1] : v;
  v = (k == 2 && n >= 2) ? ww[2] : v;
  v = (k == 3 && n >= 3) ? ww[3] : v;
  return v;
}

PXR_DEV bool esirk_axis(EnAxis &ax, int n, double x0, double dts, double v) {
  const double xo = x0 - dts * v;  // esirkepov_3d.F90:3546-3548
  const double xn = xo + dts * v;  // the one-cell loop forms the new position again (:3553-3559)
  const int first = -(n / 2);
  double wn[4] = {0.0, 0.0, 0.0, 0.0}, wo[4] = {0.0, 0.0, 0.0, 0.0};
  const int i0 = n_index(n, xn);
  w_order(n, xn - (double)i0, wn, ONESIXTH_D, TWOTHIRD_D);
  const int i1 = n_index(n, xo);
  const int d = i1 - i0;
  if (d < -1 || d > 1) return false;
  w_order(n, xo - (double)i1, wo, ONESIXTH_D, TWOTHIRD_D);
#pragma unroll
  for (int r = 0; r < 6; r++) {
    const int rel = r - 2;  // box node relative to the new node
    const double sn = pick(wn, n, rel - first);
    const double so = pick(wo, n, rel - first - d);
    ax.s0[r] = sn;
    ax.ds[r] = so - sn;
  }
  ax.base = i0;
  ax.lo = min(0, d) - n / 2;
  ax.hi = max(0, d) + (n + 1) / 2;
  return true;
}

__global__ __launch_bounds__(256) void k_esirkepov_n(EsirkNArgs a) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= a.np) return;
  const double gi = a.gi[ip];
  const double x0[3] = {(a.xp[ip] - a.xmin) * a.dxi, (a.yp[ip] - a.ymin) * a.dyi,
                        (a.zp[ip] - a.zmin) * a.dzi};
  const double v[3] = {a.ux[ip] * gi, a.uy[ip] * gi, a.uz[ip] * gi};
  EnAxis X, Y, Z;
  if (!esirk_axis(X, a.no[0], x0[0], a.dts[0], v[0]) ||
      !esirk_axis(Y, a.no[1], x0[1], a.dts[1], v[1]) ||
      !esirk_axis(Z, a.no[2], x0[2], a.dts[2], v[2]) ||
      !a.g.inside(X.base + X.lo, Y.base + Y.lo, Z.base + Z.lo) ||
      !a.g.inside(X.base + X.hi, Y.base + Y.hi, Z.base + Z.hi)) {
    count_bad(a.bad);
    return;
  }
  const double wq = a.q * a.w[ip];
  const double wqx = wq * a.invdt[0], wqy = wq * a.invdt[1], wqz = wq * a.invdt[2];
  const i64 sy = a.g.NX, sz = a.g.NX * a.g.NY;
  const i64 o = a.g.idx(X.base, Y.base, Z.base);
  // Each component's running sum runs along its own axis
  // (esirkepov_3d.F90:3811-3830); every other loop only orders the adds.
  // Jx: sum over i of wqx dsx(i) ((sy0 + dsy/2) sz0 + (sy0/2 + dsy/3) dsz)
#pragma unroll
  for (int rk = 0; rk < 6; rk++) {
    const int k = rk - 2;
    if (k < Z.lo || k > Z.hi) continue;
#pragma unroll
    for (int rj = 0; rj < 6; rj++) {
      const int j = rj - 2;
      if (j < Y.lo || j > Y.hi) continue;
      const double tw = (Y.s0[rj] + 0.5 * Y.ds[rj]) * Z.s0[rk] +
                        (0.5 * Y.s0[rj] + ONETHIRD_D * Y.ds[rj]) * Z.ds[rk];
      double run = 0.0;
      double *row = a.jx + o + sy * j + sz * k;
#pragma unroll
      for (int ri = 0; ri < 6; ri++) {
        const int i = ri - 2;
        if (i < X.lo || i >= X.hi) continue;
        const double t = wqx * X.ds[ri] * tw;
        run = (i > X.lo) ? t + run : t;
        unsafeAtomicAdd(row + i, run);
      }
    }
  }
  // Jy: running sum along j
#pragma unroll
  for (int rk = 0; rk < 6; rk++) {
    const int k = rk - 2;
    if (k < Z.lo || k > Z.hi) continue;
#pragma unroll
    for (int ri = 0; ri < 6; ri++) {
      const int i = ri - 2;
      if (i < X.lo || i > X.hi) continue;
      const double tw = (Z.s0[rk] + 0.5 * Z.ds[rk]) * X.s0[ri] +
                        (0.5 * Z.s0[rk] + ONETHIRD_D * Z.ds[rk]) * X.ds[ri];
      double run = 0.0;
      double *col = a.jy + o + i + sz * k;
#pragma unroll
      for (int rj = 0; rj < 6; rj++) {
        const int j = rj - 2;
        if (j < Y.lo || j >= Y.hi) continue;
        const double t = wqy * Y.ds[rj] * tw;
        run = (j > Y.lo) ? t + run : t;
        unsafeAtomicAdd(col + sy * j, run);
      }
    }
  }
  // Jz: running sum along k
#pragma unroll
  for (int rj = 0; rj < 6; rj++) {
    const int j = rj - 2;
    if (j < Y.lo || j > Y.hi) continue;
#pragma unroll
    for (int ri = 0; ri < 6; ri++) {
      const int i = ri - 2;
      if (i < X.lo || i > X.hi) continue;
      const double tw = (X.s0[ri] + 0.5 * X.ds[ri]) * Y.s0[rj] +
                        (0.5 * X.s0[ri] + ONETHIRD_D * X.ds[ri]) * Y.ds[rj];
      double run = 0.0;
      double *col = a.jz + o + i + sy * j;
#pragma unroll
      for (int rk = 0; rk < 6; rk++) {
        const int k = rk - 2;
        if (k < Z.lo || k >= Z.hi) continue;
        const double t = wqz * Z.ds[rk] * tw;
        run = (k > Z.lo) ? t + run : t;
        unsafeAtomicAdd(col + sz * k, run);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// charge density, pxr_depose_rho_n
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rho_n(RhoNArgs a) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= a.np) return;
  double p[3] = {(a.xp[ip] - a.xmin) * a.dxi, (a.yp[ip] - a.ymin) * a.dyi,
                 (a.zp[ip] - a.zmin) * a.dzi};
  if (a.l4symtry) {
    p[0] = fabs(p[0]);
    p[1] = fabs(p[1]);
  }
  double s[3][4];
  int first[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int j = n_index(a.no[d], p[d]);
    w_order(a.no[d], p[d] - (double)j, s[d], ONESIXTH_D, TWOTHIRD_D);
    first[d] = j - a.no[d] / 2;
  }
  if (!a.g.inside(first[0], first[1], first[2]) ||
      !a.g.inside(first[0] + a.no[0], first[1] + a.no[1], first[2] + a.no[2])) {
    count_bad(a.bad);
    return;
  }
  const double wq = a.per_particle_weight ? a.q * a.w[ip] * a.invvol : a.q * a.invvol * a.w[0];
  double *r = a.rho + a.g.idx(first[0], first[1], first[2]);
  const i64 sy = a.g.NX, sz = a.g.NX * a.g.NY;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    if (c > a.no[2]) continue;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      if (b > a.no[1]) continue;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        if (e > a.no[0]) continue;
        unsafeAtomicAdd(r + e + sy * b + sz * c, s[0][e] * s[1][b] * s[2][c] * wq);
      }
    }
  }
}

void launch_gather_n(const GatherNArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_n, dim3(grid1d(a.np, 256)), dim3(256), 0, st, a);
}

void launch_esirkepov_n(const EsirkNArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_esirkepov_n, dim3(grid1d(a.np, 256)), dim3(256), 0, st, a);
}

void launch_rho_n(const RhoNArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_rho_n, dim3(grid1d(a.np, 256)), dim3(256), 0, st, a);
}

}  // namespace pxr
