// pxr_leaf.hip -- the reference's leaf kernels on gfx950, one particle (or
// one grid node) per lane, operating directly on caller-owned device arrays.
//
// These are the drop-in counterparts of the Forthon-wrapped routines:
// geteb3d_energy_conserving, pxr_boris_push_u_3d, pxr_ebcancelpush3d,
// pxr_pushxyz, pxr_set_gamma, depose_jxjyjz, pxrpush_em3d_{e,b}vec[_norder],
// pxr_particle_bin_sorting and the single-rank periodic guard routines.
// All FP64, compiled with -ffp-contract=off: the pushers and every index
// computation follow the reference's operation order bit for bit.
#include <stdio.h>

#include "pxr_device.h"
#include "pxr_generic.h"
#include "pxr_kernels.h"
#include "pxr_leafbin.h"
#include "pxr_runtime.h"

using namespace pxr;

namespace pxr {

// ---------------------------------------------------------------------------
// field gathering
// ---------------------------------------------------------------------------
struct GatherArgs {
  i64 np;
  const double *xp, *yp, *zp;
  double *ex, *ey, *ez, *bx, *by, *bz;
  double xmin, ymin, zmin, dxi, dyi, dzi, s;
  GridGeom g;
  const double *exg, *eyg, *ezg, *bxg, *byg, *bzg;
  unsigned long long *bad;
};

template <int N, int M>
__global__ __launch_bounds__(256) void k_gather(GatherArgs a) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= a.np) return;
  const double x = (a.xp[ip] - a.xmin) * a.dxi;
  const double y = (a.yp[ip] - a.ymin) * a.dyi;
  const double z = (a.zp[ip] - a.zmin) * a.dzi;
  double out[6];
  FieldView<double> v[6];
  const double *F[6] = {a.exg, a.eyg, a.ezg, a.bxg, a.byg, a.bzg};
  for (int c = 0; c < 6; c++) v[c] = FieldView<double>{F[c] + a.g.idx(0, 0, 0), a.g.NX, a.g.NX * a.g.NY};
  if (!gather_eb<N, M, double>(x, y, z, a.s, a.g, v, out)) {
    count_bad(a.bad);
    return;
  }
  a.ex[ip] += out[0];
  a.ey[ip] += out[1];
  a.ez[ip] += out[2];
  a.bx[ip] += out[3];
  a.by[ip] += out[4];
  a.bz[ip] += out[5];
}

// ---------------------------------------------------------------------------
// pushers
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_boris(i64 np, double *ux, double *uy, double *uz,
                                               double *gi, const double *ex, const double *ey,
                                               const double *ez, const double *bx,
                                               const double *by, const double *bz, double cst) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  double u[3] = {ux[ip], uy[ip], uz[ip]};
  double g;
  boris_push<double>(u, g, ex[ip], ey[ip], ez[ip], bx[ip], by[ip], bz[ip], cst);
  ux[ip] = u[0];
  uy[ip] = u[1];
  uz[ip] = u[2];
  gi[ip] = g;
}

__global__ __launch_bounds__(256) void k_vay(i64 np, double *ux, double *uy, double *uz,
                                             double *gi, const double *ex, const double *ey,
                                             const double *ez, const double *bx, const double *by,
                                             const double *bz, double q, double m, double dt,
                                             int which) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  double u[3] = {ux[ip], uy[ip], uz[ip]};
  double g = gi[ip];
  if (which == 0) {
    vay_push<double>(u, g, ex[ip], ey[ip], ez[ip], bx[ip], by[ip], bz[ip], q * dt / m);
  } else if (which == 1) {
    vay_half1<double>(u, g, ex[ip], ey[ip], ez[ip], bx[ip], by[ip], bz[ip], 0.5 * q * dt / m);
  } else {
    vay_half2<double>(u, g, ex[ip], ey[ip], ez[ip], bx[ip], by[ip], bz[ip], 0.5 * q * dt / m);
  }
  ux[ip] = u[0];
  uy[ip] = u[1];
  uz[ip] = u[2];
  gi[ip] = g;
}

__global__ __launch_bounds__(256) void k_pushxyz(i64 np, double *xp, double *yp, double *zp,
                                                 const double *ux, const double *uy,
                                                 const double *uz, const double *gi, double dt) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  xp[ip] = xp[ip] + ux[ip] * gi[ip] * dt;
  yp[ip] = yp[ip] + uy[ip] * gi[ip] * dt;
  zp[ip] = zp[ip] + uz[ip] * gi[ip] * dt;
}

// ---------------------------------------------------------------------------
// laser antenna (laser_pusher_manager_3d.F90)
// ---------------------------------------------------------------------------
// Complex numbers as (re, im) pairs, evaluated in the reference's order of
// operations; the division is Smith's (1962) algorithm.
struct Cx {
  double re, im;
};
PXR_DEV Cx cx_mul(Cx a, Cx b) { return Cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
PXR_DEV Cx cx_scale(Cx a, double s) { return Cx{a.re * s, a.im * s}; }
PXR_DEV Cx cx_sub(Cx a, Cx b) { return Cx{a.re - b.re, a.im - b.im}; }
PXR_DEV Cx cx_add(Cx a, Cx b) { return Cx{a.re + b.re, a.im + b.im}; }
PXR_DEV Cx cx_div(Cx a, Cx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double t = b.im / b.re, den = b.re + b.im * t;
    return Cx{(a.re + a.im * t) / den, (a.im - a.re * t) / den};
  }
  const double t = b.re / b.im, den = b.im + b.re * t;
  return Cx{(a.re * t + a.im) / den, (a.im * t - a.re) / den};
}
PXR_DEV Cx cx_exp(Cx a) {
  const double e = exp(a.re);
  double sn, cs;
  sincos(a.im, &sn, &cs);
  return Cx{e * cs, e * sn};
}
// x**n, n >= 0, by binary powering (how a real**integer power is expanded)
PXR_DEV double ipow(double x, i64 n) {
  double r = 1.0;
  while (n > 0) {
    if (n & 1) r *= x;
    x *= x;
    n >>= 1;
  }
  return r;
}

struct LaserArgs {
  i64 np;
  const double *xx, *yy;  // antenna-frame coordinates (pid columns 2 and 3)
  double *x, *y, *z, *ux, *uy, *uz, *gi;
  double dt, coeff, charge;
  double emax1, emax2, pv1[3], pv2[3], k0;
  Cx qz;
  double tau, t, tpeak, polangle;
  i64 torder;
};

constexpr double PXR_PI_F90 = 3.14159265358979323;  // constants module value of pi

// WINDOW 0: laserp_pusher_gaussian + gaussian_profile (:195-271, :368-415);
// WINDOW 1: laserp_pusher_hanning + hanning_profile (:281-356, :426-478).
// clight = 1.  u = charge * amp * coeff, gaminv = 1, x += dt * u.
template <int WINDOW>
__global__ __launch_bounds__(256) void k_laser_push(LaserArgs a) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= a.np) return;
  const double xx = a.xx[ip], yy = a.yy[ip];
  const Cx j = {0.0, 1.0};
  const Cx jk = cx_scale(j, a.k0);
  double amp[3];
  if (WINDOW == 0) {
    const Cx ph = cx_scale(cx_scale(jk, 1.0), a.t - a.tpeak);
    const Cx b = cx_div(cx_scale(jk, xx * xx + yy * yy), cx_scale(a.qz, 2.0));
    Cx u1, u2;
    if (a.torder == 0) {
      u1 = cx_sub(ph, b);
      u2 = u1;
    } else {
      const Cx e = {ipow((a.t - a.tpeak) / a.tau, a.torder), 0.0};
      u1 = cx_sub(cx_sub(ph, b), e);
      u2 = cx_add(cx_sub(cx_sub(ph, b), e), cx_scale(j, a.polangle * 2.0 * PXR_PI_F90));
    }
    u1 = cx_scale(cx_exp(u1), a.emax1);
    u2 = cx_scale(cx_exp(u2), a.emax2);
    for (int d = 0; d < 3; d++) amp[d] = cx_add(cx_scale(u1, a.pv1[d]), cx_scale(u2, a.pv2[d])).im;
  } else {
    if (!(a.t <= 2.0 * a.tpeak)) {
      amp[0] = amp[1] = amp[2] = 0.0;
    } else {
      const Cx x = cx_div(Cx{xx * xx + yy * yy, 0.0}, cx_scale(a.qz, 2.0));
      const Cx arg = cx_sub(x, Cx{1.0 * (a.t - a.tpeak), 0.0});
      Cx u1 = cx_scale(cx_mul(jk, arg), -1.0);
      Cx u2 = cx_add(cx_scale(cx_mul(jk, arg), -1.0),
                     cx_scale(cx_scale(cx_scale(j, a.polangle), 2.0), PXR_PI_F90));
      const double win = 0.5 - 0.5 * cos(a.t * 2.0 * PXR_PI_F90 / (2.0 * a.tpeak));
      u1 = cx_scale(cx_scale(cx_scale(cx_exp(u1), 1.0), a.emax1), win);
      u2 = cx_scale(cx_scale(cx_scale(cx_exp(u2), 1.0), a.emax2), win);
      for (int d = 0; d < 3; d++) amp[d] = cx_add(cx_scale(u1, a.pv1[d]), cx_scale(u2, a.pv2[d])).re;
    }
  }
  const double ux = a.charge * amp[0] * a.coeff;
  const double uy = a.charge * amp[1] * a.coeff;
  const double uz = a.charge * amp[2] * a.coeff;
  a.ux[ip] = ux;
  a.uy[ip] = uy;
  a.uz[ip] = uz;
  a.gi[ip] = 1.0;
  a.x[ip] = a.x[ip] + a.dt * ux;
  a.y[ip] = a.y[ip] + a.dt * uy;
  a.z[ip] = a.z[ip] + a.dt * uz;
}

__global__ __launch_bounds__(256) void k_set_gamma(i64 np, const double *ux, const double *uy,
                                                   const double *uz, double *gi) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  gi[ip] = gamma_inv<double>(ux[ip], uy[ip], uz[ip]);
}

// ---------------------------------------------------------------------------
// current deposition (global atomics)
// ---------------------------------------------------------------------------

// Particles of the deposition kernels: a block takes 256 * DEP_S of them
// and the lanes of a wave take particles 4 * DEP_S apart (each lane DEP_S
// consecutive ones in turn).  In cell-sorted input, as the tiled species
// hold it, the lanes of one atomic instruction then sit in different cells
// instead of queueing on the same nodes.
#define DEP_S 4
PXR_DEV i64 dep_particle(int k) {
  const int t = (int)threadIdx.x;
  return (i64)blockIdx.x * (256 * DEP_S) + (t & 63) * (4 * DEP_S) + (t >> 6) * DEP_S + k;
}

template <int N>
PXR_DEV void esirkepov_one(const DepoArgs &a, i64 ip) {
  EsirkepovIn<double> in;
  esirkepov_prepare<N, double>(in, a.xp[ip], a.yp[ip], a.zp[ip], a.ux[ip], a.uy[ip], a.uz[ip],
                               a.gi[ip], a.w[ip], a.ec);
  int lo[3], hi[3];
  if (!esirkepov_box<N, double>(in, lo, hi) || !a.g.inside(lo[0], lo[1], lo[2]) ||
      !a.g.inside(hi[0], hi[1], hi[2])) {
    count_bad(a.bad);
    return;
  }
  const i64 o = a.g.idx(0, 0, 0);
  GlobalAcc acc{{a.jx + o, a.jy + o, a.jz + o}, a.g.NX, a.g.NX * a.g.NY};
  esirkepov_deposit<N, double>(in, acc);
}

template <int N>
__global__ __launch_bounds__(256) void k_esirkepov(DepoArgs a) {
  for (int k = 0; k < DEP_S; k++) {
    const i64 ip = dep_particle(k);
    if (ip < a.np) esirkepov_one<N>(a, ip);
  }
}

template <int N>
PXR_DEV void direct_one(const DepoArgs &a, i64 ip) {
  DirectIn<double> in;
  direct_prepare<N, double>(in, a.xp[ip], a.yp[ip], a.zp[ip], a.ux[ip], a.uy[ip], a.uz[ip],
                            a.gi[ip], a.w[ip], a.q, a.xmin, a.ymin, a.zmin, a.dt, a.dx, a.dy,
                            a.dz, a.s);
  int lo[3], hi[3];
  direct_box<N>(in, lo, hi);
  if (!a.g.inside(lo[0], lo[1], lo[2]) || !a.g.inside(hi[0], hi[1], hi[2])) {
    count_bad(a.bad);
    return;
  }
  const i64 o = a.g.idx(0, 0, 0);
  GlobalAcc acc{{a.jx + o, a.jy + o, a.jz + o}, a.g.NX, a.g.NX * a.g.NY};
  direct_deposit<N, double>(in, acc);
}

template <int N>
__global__ __launch_bounds__(256) void k_direct(DepoArgs a) {
  for (int k = 0; k < DEP_S; k++) {
    const i64 ip = dep_particle(k);
    if (ip < a.np) direct_one<N>(a, ip);
  }
}

// charge density, pxr_depose_rho_n (charge_deposition_3d.F90:2081-2228):
// nodal shape of order N at floor (N odd) / nint (N even), wq = q w / vol,
// rho(j+a, k+b, l+c) += sx(a) sy(b) sz(c) wq

template <int N>
PXR_DEV void rho_one(const RhoArgs &a, i64 ip) {
  double x = (a.xp[ip] - a.xmin) * a.dxi;
  double y = (a.yp[ip] - a.ymin) * a.dyi;
  const double z = (a.zp[ip] - a.zmin) * a.dzi;
  if (a.l4symtry) {
    x = fabs(x);
    y = fabs(y);
  }
  double sx[N + 1], sy[N + 1], sz[N + 1];
  const int j = shape_nodal<N, double>(x, sx);
  const int k = shape_nodal<N, double>(y, sy);
  const int l = shape_nodal<N, double>(z, sz);
  if (!a.g.inside(j, k, l) || !a.g.inside(j + N, k + N, l + N)) {
    count_bad(a.bad);
    return;
  }
  const double wq = a.per_particle_weight ? a.q * a.w[ip] * a.invvol : a.q * a.invvol * a.w[0];
  double *r = a.rho + a.g.idx(j, k, l);
#pragma unroll
  for (int c = 0; c <= N; c++)
#pragma unroll
    for (int b = 0; b <= N; b++)
#pragma unroll
      for (int e = 0; e <= N; e++)
        unsafeAtomicAdd(r + e + a.g.NX * (b + (i64)a.g.NY * c), sx[e] * sy[b] * sz[c] * wq);
}

template <int N>
__global__ __launch_bounds__(256) void k_rho(RhoArgs a) {
  for (int k = 0; k < DEP_S; k++) {
    const i64 ip = dep_particle(k);
    if (ip < a.np) rho_one<N>(a, ip);
  }
}

// ---------------------------------------------------------------------------
// FDTD
// ---------------------------------------------------------------------------
struct Arr3 {
  double *p;  // base pointer of the allocation
  int lo0, lo1, lo2;
  i64 sy, sz;
  PXR_DEV double &at(int i, int j, int k) const {
    return p[(i - lo0) + sy * (j - lo1) + sz * (k - lo2)];
  }
};

struct Box {
  int lo[3], hi[3];
};

struct YeeArgs {
  Box bx[3];   // loop box of each component
  Box un;      // union of the three
  Arr3 f[9];   // ex ey ez bx by bz [jx jy jz]
  double mudt, cx, cy, cz;
};

__global__ __launch_bounds__(256) void k_yee_e(YeeArgs a) {
  const int nxu = a.un.hi[0] - a.un.lo[0] + 1;
  const int nyu = a.un.hi[1] - a.un.lo[1] + 1;
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const i64 tot = (i64)nxu * nyu * (a.un.hi[2] - a.un.lo[2] + 1);
  if (t >= tot) return;
  const int j = a.un.lo[0] + (int)(t % nxu);
  const int k = a.un.lo[1] + (int)((t / nxu) % nyu);
  const int l = a.un.lo[2] + (int)(t / ((i64)nxu * nyu));
  const Arr3 &Ex = a.f[0], &Ey = a.f[1], &Ez = a.f[2], &Bx = a.f[3], &By = a.f[4],
             &Bz = a.f[5], &Jx = a.f[6], &Jy = a.f[7], &Jz = a.f[8];
  auto in = [&](int c) {
    return j >= a.bx[c].lo[0] && j <= a.bx[c].hi[0] && k >= a.bx[c].lo[1] &&
           k <= a.bx[c].hi[1] && l >= a.bx[c].lo[2] && l <= a.bx[c].hi[2];
  };
  // yee.F90:680-688, 698-706, 716-724
  if (in(0))
    Ex.at(j, k, l) = Ex.at(j, k, l) + a.cy * (Bz.at(j, k, l) - Bz.at(j, k - 1, l)) -
                     a.cz * (By.at(j, k, l) - By.at(j, k, l - 1)) - a.mudt * Jx.at(j, k, l);
  if (in(1))
    Ey.at(j, k, l) = Ey.at(j, k, l) - a.cx * (Bz.at(j, k, l) - Bz.at(j - 1, k, l)) +
                     a.cz * (Bx.at(j, k, l) - Bx.at(j, k, l - 1)) - a.mudt * Jy.at(j, k, l);
  if (in(2))
    Ez.at(j, k, l) = Ez.at(j, k, l) + a.cx * (By.at(j, k, l) - By.at(j - 1, k, l)) -
                     a.cy * (Bx.at(j, k, l) - Bx.at(j, k - 1, l)) - a.mudt * Jz.at(j, k, l);
}

__global__ __launch_bounds__(256) void k_yee_b(YeeArgs a) {
  const int nxu = a.un.hi[0] - a.un.lo[0] + 1;
  const int nyu = a.un.hi[1] - a.un.lo[1] + 1;
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const i64 tot = (i64)nxu * nyu * (a.un.hi[2] - a.un.lo[2] + 1);
  if (t >= tot) return;
  const int j = a.un.lo[0] + (int)(t % nxu);
  const int k = a.un.lo[1] + (int)((t / nxu) % nyu);
  const int l = a.un.lo[2] + (int)(t / ((i64)nxu * nyu));
  const Arr3 &Ex = a.f[0], &Ey = a.f[1], &Ez = a.f[2], &Bx = a.f[3], &By = a.f[4],
             &Bz = a.f[5];
  auto in = [&](int c) {
    return j >= a.bx[c].lo[0] && j <= a.bx[c].hi[0] && k >= a.bx[c].lo[1] &&
           k <= a.bx[c].hi[1] && l >= a.bx[c].lo[2] && l <= a.bx[c].hi[2];
  };
  // yee.F90:1305-1313, 1322-1330, 1339-1347
  if (in(0))
    Bx.at(j, k, l) = Bx.at(j, k, l) - a.cy * (Ez.at(j, k + 1, l) - Ez.at(j, k, l)) +
                     a.cz * (Ey.at(j, k, l + 1) - Ey.at(j, k, l));
  if (in(1))
    By.at(j, k, l) = By.at(j, k, l) + a.cx * (Ez.at(j + 1, k, l) - Ez.at(j, k, l)) -
                     a.cz * (Ex.at(j, k, l + 1) - Ex.at(j, k, l));
  if (in(2))
    Bz.at(j, k, l) = Bz.at(j, k, l) - a.cx * (Ey.at(j + 1, k, l) - Ey.at(j, k, l)) +
                     a.cy * (Ex.at(j, k + 1, l) - Ex.at(j, k, l));
}

#define PXR_MAX_HALF_ORDER 16
struct Coefs {
  double c[3][PXR_MAX_HALF_ORDER];
  int nh[3];  // norder/2 per axis
};

struct NorderArgs {
  GridGeom g;
  double *f[9];
  double mudt;
  Coefs co;
  int nxs, nys, nzs, ist;
};

PXR_DEV int imin3(int a, int b, int c) { return min(min(a, b), c); }

// Arbitrary-order FDTD, yee.F90:46-133 (E) and 752-833 (B), with stencils
// cut at the guard region.
// v +/- sum_i cf[i-1] * (f[c + (i - a) s] - f[c + (b - i) s]), term by term
// in the reference's order.  NH > 0: the full stencil of NH terms, unrolled;
// NH = 0: the n terms of a stencil cut at the guards.
template <int NH, bool PLUS>
PXR_DEV double fd_sum(double v, const double *cf, const double *f, i64 c, i64 s, int a, int b,
                      int n) {
  if (NH > 0) {
#pragma unroll
    for (int i = 1; i <= NH; i++) {
      const double t = cf[i - 1] * (f[c + (i - a) * s] - f[c + (b - i) * s]);
      v = PLUS ? v + t : v - t;
    }
  } else {
    for (int i = 1; i <= n; i++) {
      const double t = cf[i - 1] * (f[c + (i - a) * s] - f[c + (b - i) * s]);
      v = PLUS ? v + t : v - t;
    }
  }
  return v;
}

// node (j, k, l) of the loop box -s..n+s, or false past its end
PXR_DEV bool norder_node(const NorderArgs &a, int &j, int &k, int &l) {
  const GridGeom &g = a.g;
  const int nxu = g.nx + 2 * a.nxs + 1, nyu = g.ny + 2 * a.nys + 1;
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (i64)nxu * nyu * (g.nz + 2 * a.nzs + 1)) return false;
  j = -a.nxs + (int)(t % nxu);
  k = -a.nys + (int)((t / nxu) % nyu);
  l = -a.nzs + (int)(t / ((i64)nxu * nyu));
  return true;
}

// yee.F90:46-133 at one node, with nxh..nzh terms per axis
PXR_DEV void norder_e_at(const NorderArgs &a, i64 c, int nxh, int nyh, int nzh) {
  double *ex = a.f[0], *ey = a.f[1], *ez = a.f[2];
  const double *bx = a.f[3], *by = a.f[4], *bz = a.f[5], *jx = a.f[6], *jy = a.f[7], *jz = a.f[8];
  const i64 sy = a.g.NX, sz = a.g.NX * a.g.NY;
  const int ist = a.ist;
  double v = ex[c] - a.mudt * jx[c];
  v = fd_sum<0, true>(v, a.co.c[1], bz, c, sy, ist, 0, nyh);
  v = fd_sum<0, false>(v, a.co.c[2], by, c, sz, ist, 0, nzh);
  ex[c] = v;
  v = ey[c] - a.mudt * jy[c];
  v = fd_sum<0, false>(v, a.co.c[0], bz, c, 1, ist, 0, nxh);
  v = fd_sum<0, true>(v, a.co.c[2], bx, c, sz, ist, 0, nzh);
  ey[c] = v;
  v = ez[c] - a.mudt * jz[c];
  v = fd_sum<0, true>(v, a.co.c[0], by, c, 1, ist, 0, nxh);
  v = fd_sum<0, false>(v, a.co.c[1], bx, c, sy, ist, 0, nyh);
  ez[c] = v;
}

// yee.F90:752-833 at one node
PXR_DEV void norder_b_at(const NorderArgs &a, i64 c, int nxh, int nyh, int nzh) {
  const double *ex = a.f[0], *ey = a.f[1], *ez = a.f[2];
  double *bx = a.f[3], *by = a.f[4], *bz = a.f[5];
  const i64 sy = a.g.NX, sz = a.g.NX * a.g.NY;
  const int ist = a.ist;
  double v = bx[c];
  v = fd_sum<0, false>(v, a.co.c[1], ez, c, sy, 0, ist, nyh);
  v = fd_sum<0, true>(v, a.co.c[2], ey, c, sz, 0, ist, nzh);
  bx[c] = v;
  v = by[c];
  v = fd_sum<0, true>(v, a.co.c[0], ez, c, 1, 0, ist, nxh);
  v = fd_sum<0, false>(v, a.co.c[2], ex, c, sz, 0, ist, nzh);
  by[c] = v;
  v = bz[c];
  v = fd_sum<0, false>(v, a.co.c[0], ey, c, 1, 0, ist, nxh);
  v = fd_sum<0, true>(v, a.co.c[1], ex, c, sy, 0, ist, nyh);
  bz[c] = v;
}

// One thread per node, for stencils of different orders per axis or of
// more than 4 terms.
template <bool E>
__global__ __launch_bounds__(256) void k_norder(NorderArgs a) {
  int j, k, l;
  if (!norder_node(a, j, k, l)) return;
  const GridGeom &g = a.g;
  const i64 c = g.idx(j, k, l);
  const int nxh = imin3(a.co.nh[0], (g.nx - j) + g.gx, j + g.gx);
  const int nyh = imin3(a.co.nh[1], (g.ny - k) + g.gy, k + g.gy);
  const int nzh = imin3(a.co.nh[2], (g.nz - l) + g.gz, l + g.gz);
  if (E) norder_e_at(a, c, nxh, nyh, nzh);
  else norder_b_at(a, c, nxh, nyh, nzh);
}

// Register-blocked form for a full stencil of NH terms on every axis and a
// staggering shift IST known at compile time.  A thread owns one (x, y)
// column of the loop box and marches up ZCHUNK planes of it.  The two fields
// differenced along z (B_y and B_x for E; E_y and E_x for B) are kept as a
// window of 2 NH + 1 planes in registers, so each plane of them is loaded
// once; the x and y differences read the plane from memory (neighbouring
// lanes and the block's rows share those lines in L1/L2).  Nodes whose
// stencil is cut at the guards take the per-node form.  The terms are
// summed in the reference's order, so the results equal k_norder's.
#ifndef ZCHUNK
#define ZCHUNK 64  // planes per thread
#endif
#ifndef ZBX
#define ZBX 64     // columns of a block along x (256 / ZBX along y)
#endif
template <int NH, bool E, int IST>
__global__ __launch_bounds__(256) void k_norder_z(NorderArgs a) {
  const GridGeom &g = a.g;
  const int nxu = g.nx + 2 * a.nxs + 1, nyu = g.ny + 2 * a.nys + 1, nzu = g.nz + 2 * a.nzs + 1;
  const int ju = (int)blockIdx.x * ZBX + (int)(threadIdx.x % ZBX);
  const int ku = (int)blockIdx.y * (256 / ZBX) + (int)(threadIdx.x / ZBX);
  if (ju >= nxu || ku >= nyu) return;
  const int j = -a.nxs + ju, k = -a.nys + ku;
  const int l0 = -a.nzs + (int)blockIdx.z * ZCHUNK;
  const int l1 = min(l0 + ZCHUNK, -a.nzs + nzu);  // exclusive
  const int nxh = imin3(a.co.nh[0], (g.nx - j) + g.gx, j + g.gx);
  const int nyh = imin3(a.co.nh[1], (g.ny - k) + g.gy, k + g.gy);
  // planes whose z stencil is whole: -gz + NH .. nz + gz - NH
  int m0 = max(l0, -g.gz + NH), m1 = min(l1, g.nz + g.gz - NH + 1);
  if (nxh != NH || nyh != NH || m0 >= m1) m0 = m1 = l1;  // no marching part
  auto node = [&](int l) {
    const int nzh = imin3(a.co.nh[2], (g.nz - l) + g.gz, l + g.gz);
    if (E) norder_e_at(a, g.idx(j, k, l), nxh, nyh, nzh);
    else norder_b_at(a, g.idx(j, k, l), nxh, nyh, nzh);
  };
  for (int l = l0; l < m0; l++) node(l);
  if (m0 < m1) {
    const i64 sy = g.NX, sz = g.NX * g.NY;
    const double *c0 = a.co.c[0], *c1 = a.co.c[1], *c2 = a.co.c[2];
    // u, w: the two fields differenced along z, planes l - NH .. l + NH
    const double *fu = E ? a.f[4] : a.f[1];  // B_y | E_y
    const double *fw = E ? a.f[3] : a.f[0];  // B_x | E_x
    double u[2 * NH + 1], w[2 * NH + 1];
    i64 c = g.idx(j, k, m0);
#pragma unroll
    for (int o = 0; o < 2 * NH; o++) {
      u[o] = fu[c + (o - NH) * sz];
      w[o] = fw[c + (o - NH) * sz];
    }
    for (int l = m0; l < m1; l++, c += sz) {
      u[2 * NH] = fu[c + NH * sz];
      w[2 * NH] = fw[c + NH * sz];
      if (E) {
        double *ex = a.f[0], *ey = a.f[1], *ez = a.f[2];
        const double *bx = a.f[3], *by = a.f[4], *bz = a.f[5];
        const double *jx = a.f[6], *jy = a.f[7], *jz = a.f[8];
        double v = ex[c] - a.mudt * jx[c];
        v = fd_sum<NH, true>(v, c1, bz, c, sy, IST, 0, NH);
#pragma unroll
        for (int i = 1; i <= NH; i++) v = v - c2[i - 1] * (u[NH + i - IST] - u[NH - i]);
        ex[c] = v;
        v = ey[c] - a.mudt * jy[c];
        v = fd_sum<NH, false>(v, c0, bz, c, 1, IST, 0, NH);
#pragma unroll
        for (int i = 1; i <= NH; i++) v = v + c2[i - 1] * (w[NH + i - IST] - w[NH - i]);
        ey[c] = v;
        v = ez[c] - a.mudt * jz[c];
        v = fd_sum<NH, true>(v, c0, by, c, 1, IST, 0, NH);
        v = fd_sum<NH, false>(v, c1, bx, c, sy, IST, 0, NH);
        ez[c] = v;
      } else {
        const double *ex = a.f[0], *ey = a.f[1], *ez = a.f[2];
        double *bx = a.f[3], *by = a.f[4], *bz = a.f[5];
        double v = bx[c];
        v = fd_sum<NH, false>(v, c1, ez, c, sy, 0, IST, NH);
#pragma unroll
        for (int i = 1; i <= NH; i++) v = v + c2[i - 1] * (u[NH + i] - u[NH - i + IST]);
        bx[c] = v;
        v = by[c];
        v = fd_sum<NH, true>(v, c0, ez, c, 1, 0, IST, NH);
#pragma unroll
        for (int i = 1; i <= NH; i++) v = v - c2[i - 1] * (w[NH + i] - w[NH - i + IST]);
        by[c] = v;
        v = bz[c];
        v = fd_sum<NH, false>(v, c0, ey, c, 1, 0, IST, NH);
        v = fd_sum<NH, true>(v, c1, ex, c, sy, 0, IST, NH);
        bz[c] = v;
      }
#pragma unroll
      for (int o = 0; o < 2 * NH; o++) {
        u[o] = u[o + 1];
        w[o] = w[o + 1];
      }
    }
  }
  for (int l = max(m1, l0); l < l1; l++) node(l);
}

template <int NH, bool E>
static void launch_norder_z(const NorderArgs &a, hipStream_t st) {
  const int nxu = a.g.nx + 2 * a.nxs + 1, nyu = a.g.ny + 2 * a.nys + 1;
  const int nzu = a.g.nz + 2 * a.nzs + 1;
  const dim3 gr((nxu + ZBX - 1) / ZBX, (nyu + 256 / ZBX - 1) / (256 / ZBX),
               (nzu + ZCHUNK - 1) / ZCHUNK), bl(256);
  if (a.ist) hipLaunchKernelGGL((k_norder_z<NH, E, 1>), gr, bl, 0, st, a);
  else hipLaunchKernelGGL((k_norder_z<NH, E, 0>), gr, bl, 0, st, a);
}

// the register-blocked kernel for a stencil of nh <= 4 terms on every axis,
// the per-node one otherwise
static void launch_norder(const NorderArgs &a, bool e, i64 tot) {
  const int nh = a.co.nh[0];
  const int u = (nh == a.co.nh[1] && nh == a.co.nh[2] && nh <= 4) ? nh : 0;
  hipStream_t st = rt().stream;
  switch (u * 2 + (e ? 1 : 0)) {
    case 2: launch_norder_z<1, false>(a, st); break;
    case 3: launch_norder_z<1, true>(a, st); break;
    case 4: launch_norder_z<2, false>(a, st); break;
    case 5: launch_norder_z<2, true>(a, st); break;
    case 6: launch_norder_z<3, false>(a, st); break;
    case 7: launch_norder_z<3, true>(a, st); break;
    case 8: launch_norder_z<4, false>(a, st); break;
    case 9: launch_norder_z<4, true>(a, st); break;
    default: {
      const dim3 gr(grid1d(tot, 256)), bl(256);
      if (e) hipLaunchKernelGGL((k_norder<true>), gr, bl, 0, st, a);
      else hipLaunchKernelGGL((k_norder<false>), gr, bl, 0, st, a);
    }
  }
}

// ---------------------------------------------------------------------------
// periodic guards on one rank
// ---------------------------------------------------------------------------
// One axis of field_bc: nodes n..n+g take nodes 0..g, nodes -g..-1 take
// n-g..n-1 (node 0 is unchanged: it receives node n, itself a copy of 0).
__global__ __launch_bounds__(256) void k_field_bc_axis(double *f, GridGeom g, int axis) {
  int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  int gg = axis == 0 ? g.gx : axis == 1 ? g.gy : g.gz;
  const int e0 = g.nx + 1 + 2 * g.gx, e1 = g.ny + 1 + 2 * g.gy, e2 = g.nz + 1 + 2 * g.gz;
  int ext[3] = {e0, e1, e2};
  ext[axis] = 2 * gg + 1;  // gg nodes below, gg+1 nodes from n up
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (i64)ext[0] * ext[1] * ext[2]) return;
  int c[3];
  c[0] = (int)(t % ext[0]);
  c[1] = (int)((t / ext[0]) % ext[1]);
  c[2] = (int)(t / ((i64)ext[0] * ext[1]));
  int lo[3] = {-g.gx, -g.gy, -g.gz};
  for (int d = 0; d < 3; d++) c[d] += lo[d];
  // along the axis: r in [-gg, gg]; r < 0 -> node r (lower guard) gets r+n;
  // r >= 0 -> node n+r gets r
  int r = c[axis];
  int dst[3] = {c[0], c[1], c[2]}, src[3] = {c[0], c[1], c[2]};
  if (r < 0) {
    dst[axis] = r;
    src[axis] = r + n;
  } else {
    dst[axis] = n + r;
    src[axis] = r;
  }
  f[g.idx(dst[0], dst[1], dst[2])] = f[g.idx(src[0], src[1], src[2])];
}

// One axis of summation_bcs: a(0:g) += a(n:n+g); then a(n-g:n) += a(-g:0)
// with a(0) already updated.  Written per node in closed form; node 0 and
// node n are handled by one lane since each reads the other.
__global__ __launch_bounds__(256) void k_current_bc_axis(double *a, GridGeom g, int axis) {
  int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  int gg = axis == 0 ? g.gx : axis == 1 ? g.gy : g.gz;
  int ext[3] = {g.nx + 1 + 2 * g.gx, g.ny + 1 + 2 * g.gy, g.nz + 1 + 2 * g.gz};
  ext[axis] = 2 * gg + 1;  // r in [0, gg]: low side; r in [gg+1, 2gg]: high side
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (i64)ext[0] * ext[1] * ext[2]) return;
  int c[3];
  c[0] = (int)(t % ext[0]);
  c[1] = (int)((t / ext[0]) % ext[1]);
  c[2] = (int)(t / ((i64)ext[0] * ext[1]));
  int lo[3] = {-g.gx, -g.gy, -g.gz};
  for (int d = 0; d < 3; d++)
    if (d != axis) c[d] += lo[d];
  const int r = c[axis];
  int p[3] = {c[0], c[1], c[2]}, q[3] = {c[0], c[1], c[2]};
  if (r == 0) {  // node 0 += node n; node n += (updated) node 0
    p[axis] = 0;
    q[axis] = n;
    const i64 i0 = g.idx(p[0], p[1], p[2]), in = g.idx(q[0], q[1], q[2]);
    const double a0 = a[i0], an = a[in];
    const double s0 = a0 + an;
    a[i0] = s0;
    a[in] = an + s0;
  } else if (r <= gg) {  // node r (1..gg) receives node n+r
    p[axis] = r;
    q[axis] = n + r;
    const i64 ip = g.idx(p[0], p[1], p[2]);
    a[ip] = a[ip] + a[g.idx(q[0], q[1], q[2])];
  } else {  // node n-gg+s (s = r-gg-1 in 0..gg-1) receives node -gg+s
    const int s = r - gg - 1;
    p[axis] = n - gg + s;
    q[axis] = -gg + s;
    const i64 ip = g.idx(p[0], p[1], p[2]);
    a[ip] = a[ip] + a[g.idx(q[0], q[1], q[2])];
  }
}

// ---------------------------------------------------------------------------
// stable counting sort (pxr_particle_bin_sorting)
// ---------------------------------------------------------------------------
struct SortGeom {
  double xmin, ymin, zmin, dxi, dyi, dzi;
  i64 nx3, ny3, nbins;
};

__global__ __launch_bounds__(256) void k_fill_i32(int *p, i64 n, int v) {
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ __launch_bounds__(64) void k_fill_u64(unsigned long long *p, i64 n,
                                                 unsigned long long v) {
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// key (0-based bin = reference key - 1), histogram, range check
__global__ __launch_bounds__(256) void k_sort_keys(i64 np, const double *xp, const double *yp,
                                                   const double *zp, SortGeom s, int *key,
                                                   int *hist, unsigned long long *bad) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  const i64 ix = (i64)floor((xp[ip] - s.xmin) * s.dxi);
  const i64 iy = (i64)floor((yp[ip] - s.ymin) * s.dyi);
  const i64 iz = (i64)floor((zp[ip] - s.zmin) * s.dzi);
  const i64 k = iz * s.nx3 * s.ny3 + iy * s.nx3 + ix;  // reference key - 1 (sorting.F90:304)
  if (k < 0 || k >= s.nbins) {
    count_bad(bad);
    key[ip] = -1;
    return;
  }
  key[ip] = (int)k;
  atomicAdd(hist + k, 1);
}

// exclusive scan of n ints, one workgroup of 1024 lanes (bins of one sort)
__global__ __launch_bounds__(1024) void k_scan_1wg(int *v, i64 n) {
  __shared__ int s[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (i64 base = 0; base < n; base += 1024) {
    const i64 i = base + threadIdx.x;
    int x = i < n ? v[i] : 0;
    s[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int y = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += y;
      __syncthreads();
    }
    const int incl = s[threadIdx.x];
    if (i < n) v[i] = carry + incl - x;
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
}

// Stable scatter: destination of every particle, input order preserved
// among equal keys.  One workgroup walks the particles in rounds of 1024;
// inside a round each wave ranks its lanes among equal keys (ballot), and
// the waves take turns, in order, to claim their slots from the bin
// cursors, so the claim order is the input order.
__global__ __launch_bounds__(1024) void k_sort_scatter_1wg(i64 np, const int *key, int *cursor,
                                                          int *dest) {
  __shared__ int s_slot[1024];
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (i64 base = 0; base < np; base += blockDim.x) {
    const i64 ip = base + threadIdx.x;
    const bool act = ip < np;
    const int k = act ? key[ip] : -1;
    int rank, cnt, leader_lane;
    wave_peers(act, k, rank, cnt, leader_lane);
    for (int w = 0; w < nw; w++) {
      // storing the returned value makes the wave wait for its atomic
      // before the barrier, so the claims land in wave order
      if (w == wave && act && rank == 0) s_slot[threadIdx.x] = atomicAdd(cursor + k, cnt);
      __syncthreads();
    }
    if (act) dest[ip] = s_slot[(wave << 6) + leader_lane] + rank;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_permute(i64 np, const double *src, double *dst,
                                                 const int *dest) {
  const i64 ip = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip < np) dst[dest[ip]] = src[ip];
}

}  // namespace pxr

// ===========================================================================
// C ABI
// ===========================================================================
static GridGeom mkgeom(int64_t nx, int64_t ny, int64_t nz, int64_t gx, int64_t gy, int64_t gz) {
  GridGeom g;
  g.nx = (int)nx;
  g.ny = (int)ny;
  g.nz = (int)nz;
  g.gx = (int)gx;
  g.gy = (int)gy;
  g.gz = (int)gz;
  g.NX = nx + 1 + 2 * gx;
  g.NY = ny + 1 + 2 * gy;
  return g;
}

static int check_grid(int64_t nx, int64_t ny, int64_t nz, int64_t gx, int64_t gy, int64_t gz) {
  if (nx < 0 || ny < 0 || nz < 0 || gx < 0 || gy < 0 || gz < 0)
    return set_error(PXRHIP_ERR_ARG, "negative grid size or guard count");
  double tot = (double)(nx + 1 + 2 * gx) * (double)(ny + 1 + 2 * gy) * (double)(nz + 1 + 2 * gz);
  if (nx > (1 << 24) || ny > (1 << 24) || nz > (1 << 24) || tot > 9.0e15)
    return set_error(PXRHIP_ERR_ARG, "grid too large");
  return PXRHIP_OK;
}

static int finish_checked(const char *what) {
  if (!rt().checks) return PXRHIP_OK;
  unsigned long long n = 0;
  int r = take_bad(&n);
  if (r != PXRHIP_OK) return r;
  if (n)
    return set_error(PXRHIP_ERR_BOUNDS, "%s: %llu particle(s) outside the grid arrays were skipped",
                     what, n);
  return PXRHIP_OK;
}

extern "C" {

int pxrhip_geteb3d_energy_conserving(
    int64_t np, const double *xp, const double *yp, const double *zp, double *ex, double *ey,
    double *ez, double *bx, double *by, double *bz, double xmin, double ymin, double zmin,
    double dx, double dy, double dz, int64_t nx, int64_t ny, int64_t nz, int64_t nxguard,
    int64_t nyguard, int64_t nzguard, int64_t nox, int64_t noy, int64_t noz, const double *exg,
    const double *eyg, const double *ezg, const double *bxg, const double *byg, const double *bzg,
    int64_t ll4symtry, int64_t l_lower_order_in_v, int64_t lvect, int64_t field_gathe_algo,
    int64_t l_nodal) {
  (void)lvect;
  PXR_INIT();
  if (np < 0) return set_error(PXRHIP_ERR_ARG, "np < 0");
  if (np == 0) return PXRHIP_OK;  // field_gathering_manager_3d.F90:418
  int r = check_grid(nx, ny, nz, nxguard, nyguard, nzguard);
  if (r) return r;
  if (nox < 1 || nox > 3 || noy < 1 || noy > 3 || noz < 1 || noz > 3)
    return set_error(PXRHIP_ERR_ARG,
                     "shape orders (%lld,%lld,%lld): the reference's gather has weights for "
                     "orders 1..3 on each axis", (long long)nox, (long long)noy, (long long)noz);
  if (!xp || !yp || !zp || !ex || !ey || !ez || !bx || !by || !bz || !exg || !eyg || !ezg || !bxg ||
      !byg || !bzg)
    return set_error(PXRHIP_ERR_ARG, "null array");
  // geteb3d_energy_conserving_generic (field_gathering_manager_3d.F90:397-700):
  // field_gathe_algo 2 and orders other than 1_1_1, 2_2_2, 3_3_3 run the
  // generic routines, the only ones that read ll4symtry
  if (field_gathe_algo == 2 || nox != noy || nox != noz) {
    GatherNArgs n;
    n.np = np;
    n.xp = xp; n.yp = yp; n.zp = zp;
    n.ex = ex; n.ey = ey; n.ez = ez; n.bx = bx; n.by = by; n.bz = bz;
    n.xmin = xmin; n.ymin = ymin; n.zmin = zmin;
    n.dxi = 1.0 / dx; n.dyi = 1.0 / dy; n.dzi = 1.0 / dz;
    n.s = l_nodal ? 0.0 : 0.5;
    n.g = mkgeom(nx, ny, nz, nxguard, nyguard, nzguard);
    n.exg = exg; n.eyg = eyg; n.ezg = ezg; n.bxg = bxg; n.byg = byg; n.bzg = bzg;
    n.no[0] = (int)nox; n.no[1] = (int)noy; n.no[2] = (int)noz;
    n.lower = l_lower_order_in_v != 0;
    n.l4symtry = ll4symtry != 0;
    n.bad = rt().d_bad;
    launch_gather_n(n, rt().stream);
    PXR_LAUNCHED();
    return finish_checked("geteb3d_energy_conserving");
  }
  GatherArgs a;
  a.np = np;
  a.xp = xp; a.yp = yp; a.zp = zp;
  a.ex = ex; a.ey = ey; a.ez = ez; a.bx = bx; a.by = by; a.bz = bz;
  a.xmin = xmin; a.ymin = ymin; a.zmin = zmin;
  a.dxi = 1.0 / dx; a.dyi = 1.0 / dy; a.dzi = 1.0 / dz;
  a.s = l_nodal ? 0.0 : 0.5;
  a.g = mkgeom(nx, ny, nz, nxguard, nyguard, nzguard);
  a.exg = exg; a.eyg = eyg; a.ezg = ezg; a.bxg = bxg; a.byg = byg; a.bzg = bzg;
  a.bad = rt().d_bad;
  const bool lo = l_lower_order_in_v != 0;
  dim3 grd(grid1d(np, 256)), blk(256);
  hipStream_t st = rt().stream;
  switch (nox * 2 + (lo ? 1 : 0)) {
    case 2: hipLaunchKernelGGL((k_gather<1, 1>), grd, blk, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_gather<1, 0>), grd, blk, 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_gather<2, 2>), grd, blk, 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_gather<2, 1>), grd, blk, 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_gather<3, 3>), grd, blk, 0, st, a); break;
    default: hipLaunchKernelGGL((k_gather<3, 2>), grd, blk, 0, st, a); break;
  }
  PXR_LAUNCHED();
  return finish_checked("geteb3d_energy_conserving");
}

int pxrhip_boris_push_u_3d(int64_t np, double *uxp, double *uyp, double *uzp, double *gaminv,
                           const double *ex, const double *ey, const double *ez, const double *bx,
                           const double *by, const double *bz, double q, double m, double dt) {
  PXR_INIT();
  if (np < 0) return set_error(PXRHIP_ERR_ARG, "np < 0");
  if (np == 0) return PXRHIP_OK;
  const double cst = q * dt * 0.5 / m;  // boris_3d.F90:83
  hipLaunchKernelGGL(k_boris, dim3(grid1d(np, 256)), dim3(256), 0, rt().stream, np, uxp, uyp,
                     uzp, gaminv, ex, ey, ez, bx, by, bz, cst);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_ebcancelpush3d(int64_t np, double *uxp, double *uyp, double *uzp, double *gi,
                          const double *ex, const double *ey, const double *ez, const double *bx,
                          const double *by, const double *bz, double q, double m, double dt,
                          int64_t which) {
  PXR_INIT();
  if (np < 0) return set_error(PXRHIP_ERR_ARG, "np < 0");
  if (which < 0 || which > 2) return set_error(PXRHIP_ERR_ARG, "which must be 0, 1 or 2");
  if (np == 0) return PXRHIP_OK;
  hipLaunchKernelGGL(k_vay, dim3(grid1d(np, 256)), dim3(256), 0, rt().stream, np, uxp, uyp, uzp,
                     gi, ex, ey, ez, bx, by, bz, q, m, dt, (int)which);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_pushxyz(int64_t np, double *xp, double *yp, double *zp, const double *uxp,
                   const double *uyp, const double *uzp, const double *gaminv, double dt) {
  PXR_INIT();
  if (np < 0) return set_error(PXRHIP_ERR_ARG, "np < 0");
  if (np == 0) return PXRHIP_OK;
  hipLaunchKernelGGL(k_pushxyz, dim3(grid1d(np, 256)), dim3(256), 0, rt().stream, np, xp, yp, zp,
                     uxp, uyp, uzp, gaminv, dt);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

static int laser_push(int window, int64_t np, int64_t npidd, const double *pid, double *xp,
                      double *yp, double *zp, double *uxp, double *uyp, double *uzp,
                      double *gaminv, double dtt, double emax, double emax1, double emax2,
                      const double *polvector1, const double *polvector2, double k0_laser,
                      const double *q_z, double charge, double laser_tau, double real_time,
                      double t_peak, int64_t temporal_order, double polangle) {
  PXR_INIT();
  if (np < 0) return set_error(PXRHIP_ERR_ARG, "np < 0");
  if (np == 0) return PXRHIP_OK;
  if (npidd < 3) return set_error(PXRHIP_ERR_ARG, "npidd must be >= 3 (pid columns 2 and 3 hold the antenna frame)");
  if (!pid || !xp || !yp || !zp || !uxp || !uyp || !uzp || !gaminv)
    return set_error(PXRHIP_ERR_ARG, "null particle array");
  if (!polvector1 || !polvector2 || !q_z) return set_error(PXRHIP_ERR_ARG, "null laser parameter");
  if (temporal_order < 0) return set_error(PXRHIP_ERR_ARG, "temporal_order < 0");
  if (!(emax != 0.0)) return set_error(PXRHIP_ERR_ARG, "emax must be non-zero");
  LaserArgs a;
  a.np = np;
  a.xx = pid + np;
  a.yy = pid + 2 * np;
  a.x = xp; a.y = yp; a.z = zp;
  a.ux = uxp; a.uy = uyp; a.uz = uzp; a.gi = gaminv;
  a.dt = dtt;
  a.coeff = (0.01 * 1.0) / emax;  // disp_max / emax, disp_max = 0.01 c
  a.charge = charge;
  a.emax1 = emax1;
  a.emax2 = emax2;
  for (int d = 0; d < 3; d++) {
    a.pv1[d] = polvector1[d];
    a.pv2[d] = polvector2[d];
  }
  a.k0 = k0_laser;
  a.qz = Cx{q_z[0], q_z[1]};
  a.tau = laser_tau;
  a.t = real_time;
  a.tpeak = t_peak;
  a.polangle = polangle;
  a.torder = temporal_order;
  if (window == 0)
    hipLaunchKernelGGL(k_laser_push<0>, dim3(grid1d(np, 256)), dim3(256), 0, rt().stream, a);
  else
    hipLaunchKernelGGL(k_laser_push<1>, dim3(grid1d(np, 256)), dim3(256), 0, rt().stream, a);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_laserp_pusher_gaussian(int64_t np, int64_t npidd, const double *pid, double *xp,
                                  double *yp, double *zp, double *uxp, double *uyp, double *uzp,
                                  double *gaminv, double dtt, int64_t lvect, double emax,
                                  double emax1, double emax2, const double *polvector1,
                                  const double *polvector2, double k0_laser, const double *q_z,
                                  double charge, double laser_tau, double real_time,
                                  double t_peak, int64_t temporal_order, double polangle) {
  (void)lvect;
  return laser_push(0, np, npidd, pid, xp, yp, zp, uxp, uyp, uzp, gaminv, dtt, emax, emax1,
                    emax2, polvector1, polvector2, k0_laser, q_z, charge, laser_tau, real_time,
                    t_peak, temporal_order, polangle);
}

int pxrhip_laserp_pusher_hanning(int64_t np, int64_t npidd, const double *pid, double *xp,
                                 double *yp, double *zp, double *uxp, double *uyp, double *uzp,
                                 double *gaminv, double dtt, int64_t lvect, double emax,
                                 double emax1, double emax2, const double *polvector1,
                                 const double *polvector2, double k0_laser, const double *q_z,
                                 double charge, double real_time, double t_peak,
                                 int64_t temporal_order, double polangle) {
  (void)lvect;
  return laser_push(1, np, npidd, pid, xp, yp, zp, uxp, uyp, uzp, gaminv, dtt, emax, emax1,
                    emax2, polvector1, polvector2, k0_laser, q_z, charge, 1.0, real_time, t_peak,
                    temporal_order, polangle);
}

int pxrhip_set_gamma(int64_t np, const double *uxp, const double *uyp, const double *uzp,
                     double *gaminv) {
  PXR_INIT();
  if (np < 0) return set_error(PXRHIP_ERR_ARG, "np < 0");
  if (np == 0) return PXRHIP_OK;
  hipLaunchKernelGGL(k_set_gamma, dim3(grid1d(np, 256)), dim3(256), 0, rt().stream, np, uxp, uyp,
                     uzp, gaminv);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_depose_jxjyjz(double *jx, double *jy, double *jz, int64_t np, const double *xp,
                         const double *yp, const double *zp, const double *uxp, const double *uyp,
                         const double *uzp, const double *gaminv, const double *w, double q,
                         double xmin, double ymin, double zmin, double dt, double dx, double dy,
                         double dz, int64_t nx, int64_t ny, int64_t nz, int64_t nxguard,
                         int64_t nyguard, int64_t nzguard, int64_t nox, int64_t noy, int64_t noz,
                         int64_t current_depo_algo, int64_t l_nodal) {
  PXR_INIT();
  if (np < 0) return set_error(PXRHIP_ERR_ARG, "np < 0");
  if (np == 0) return PXRHIP_OK;
  int r = check_grid(nx, ny, nz, nxguard, nyguard, nzguard);
  if (r) return r;
  if (nox < 0 || nox > 3 || noy < 0 || noy > 3 || noz < 0 || noz > 3)
    return set_error(PXRHIP_ERR_ARG,
                     "shape orders (%lld,%lld,%lld): the reference's deposition has weights for "
                     "orders 0..3 on each axis", (long long)nox, (long long)noy, (long long)noz);
  if (current_depo_algo < 0 || current_depo_algo > 3)
    return set_error(PXRHIP_ERR_ARG, "current_depo_algo must be 0..3");
  if (!jx || !jy || !jz || !xp || !yp || !zp || !uxp || !uyp || !uzp || !gaminv || !w)
    return set_error(PXRHIP_ERR_ARG, "null array");
  // depose_jxjyjz_generic (current_deposition_manager_3d.F90:139-206):
  // current_depo_algo 1, and any algorithm with orders other than 1_1_1,
  // 2_2_2, 3_3_3, run pxr_depose_jxjyjz_esirkepov_n
  if (current_depo_algo == 1 || nox != noy || nox != noz || nox == 0) {
    EsirkNArgs n;
    n.np = np;
    n.xp = xp; n.yp = yp; n.zp = zp; n.ux = uxp; n.uy = uyp; n.uz = uzp; n.gi = gaminv; n.w = w;
    n.q = q; n.xmin = xmin; n.ymin = ymin; n.zmin = zmin;
    n.dxi = 1.0 / dx; n.dyi = 1.0 / dy; n.dzi = 1.0 / dz;  // esirkepov_3d.F90:3523-3534
    n.dts[0] = dt * n.dxi; n.dts[1] = dt * n.dyi; n.dts[2] = dt * n.dzi;
    n.invdt[0] = 1.0 / (dt * dy * dz);
    n.invdt[1] = 1.0 / (dt * dx * dz);
    n.invdt[2] = 1.0 / (dt * dx * dy);
    n.g = mkgeom(nx, ny, nz, nxguard, nyguard, nzguard);
    n.jx = jx; n.jy = jy; n.jz = jz;
    n.no[0] = (int)nox; n.no[1] = (int)noy; n.no[2] = (int)noz;
    n.bad = rt().d_bad;
    launch_esirkepov_n(n, rt().stream);
    PXR_LAUNCHED();
    return finish_checked("depose_jxjyjz");
  }
  DepoArgs a;
  a.np = np;
  a.xp = xp; a.yp = yp; a.zp = zp; a.ux = uxp; a.uy = uyp; a.uz = uzp; a.gi = gaminv; a.w = w;
  a.q = q; a.xmin = xmin; a.ymin = ymin; a.zmin = zmin; a.dt = dt; a.dx = dx; a.dy = dy; a.dz = dz;
  a.s = l_nodal ? 0.0 : 0.5;
  a.ec = make_esirk_const<double>(q, xmin, ymin, zmin, dt, dx, dy, dz);
  a.g = mkgeom(nx, ny, nz, nxguard, nyguard, nzguard);
  a.jx = jx; a.jy = jy; a.jz = jz;
  a.bad = rt().d_bad;
  dim3 grd(grid1d(np, 256 * DEP_S)), blk(256);
  hipStream_t st = rt().stream;
  const bool direct = current_depo_algo >= 2;  // depose_jxjyjz_generic CASE(2), CASE(3)
  if (!direct && leaf_binned_esirkepov_pays(np, a.g)) {
    r = leaf_binned_esirkepov(a, (int)nox, st);
    if (r) return r;
  } else if (direct) {
    if (nox == 1) hipLaunchKernelGGL(k_direct<1>, grd, blk, 0, st, a);
    else if (nox == 2) hipLaunchKernelGGL(k_direct<2>, grd, blk, 0, st, a);
    else hipLaunchKernelGGL(k_direct<3>, grd, blk, 0, st, a);
  } else {
    if (nox == 1) hipLaunchKernelGGL(k_esirkepov<1>, grd, blk, 0, st, a);
    else if (nox == 2) hipLaunchKernelGGL(k_esirkepov<2>, grd, blk, 0, st, a);
    else hipLaunchKernelGGL(k_esirkepov<3>, grd, blk, 0, st, a);
  }
  PXR_LAUNCHED();
  return finish_checked("depose_jxjyjz");
}

int pxrhip_depose_rho_n(double *rho, int64_t np, const double *xp, const double *yp,
                        const double *zp, const double *w, double q, double xmin, double ymin,
                        double zmin, double dx, double dy, double dz, int64_t nx, int64_t ny,
                        int64_t nz, int64_t nxguard, int64_t nyguard, int64_t nzguard, int64_t nox,
                        int64_t noy, int64_t noz, int64_t l_particles_weight, int64_t l4symtry) {
  PXR_INIT();
  if (np < 0) return set_error(PXRHIP_ERR_ARG, "np < 0");
  if (np == 0) return PXRHIP_OK;
  int r = check_grid(nx, ny, nz, nxguard, nyguard, nzguard);
  if (r) return r;
  if (!rho || !xp || !yp || !zp || !w) return set_error(PXRHIP_ERR_ARG, "null array");
  if (nox < 0 || nox > 3 || noy < 0 || noy > 3 || noz < 0 || noz > 3)
    return set_error(PXRHIP_ERR_ARG,
                     "shape orders (%lld,%lld,%lld): pxr_depose_rho_n has weights for orders "
                     "0..3 on each axis", (long long)nox, (long long)noy, (long long)noz);
  if (nox != noy || nox != noz || nox == 0) {
    RhoNArgs n;
    n.np = np;
    n.xp = xp; n.yp = yp; n.zp = zp; n.w = w;
    n.q = q; n.xmin = xmin; n.ymin = ymin; n.zmin = zmin;
    n.dxi = 1.0 / dx; n.dyi = 1.0 / dy; n.dzi = 1.0 / dz;
    n.invvol = n.dxi * n.dyi * n.dzi;
    n.per_particle_weight = l_particles_weight != 0;
    n.l4symtry = l4symtry != 0;
    n.g = mkgeom(nx, ny, nz, nxguard, nyguard, nzguard);
    n.rho = rho;
    n.no[0] = (int)nox; n.no[1] = (int)noy; n.no[2] = (int)noz;
    n.bad = rt().d_bad;
    launch_rho_n(n, rt().stream);
    PXR_LAUNCHED();
    return finish_checked("pxr_depose_rho_n");
  }
  RhoArgs a;
  a.np = np;
  a.xp = xp; a.yp = yp; a.zp = zp; a.w = w;
  a.q = q; a.xmin = xmin; a.ymin = ymin; a.zmin = zmin;
  a.dxi = 1.0 / dx; a.dyi = 1.0 / dy; a.dzi = 1.0 / dz;
  a.invvol = a.dxi * a.dyi * a.dzi;
  a.per_particle_weight = l_particles_weight != 0;
  a.l4symtry = l4symtry != 0;
  a.g = mkgeom(nx, ny, nz, nxguard, nyguard, nzguard);
  a.rho = rho;
  a.bad = rt().d_bad;
  dim3 grd(grid1d(np, 256 * DEP_S)), blk(256);
  hipStream_t st = rt().stream;
  if (leaf_binned_rho_pays(np, a.g)) {
    r = leaf_binned_rho(a, (int)nox, st);
    if (r) return r;
  } else if (nox == 1) hipLaunchKernelGGL(k_rho<1>, grd, blk, 0, st, a);
  else if (nox == 2) hipLaunchKernelGGL(k_rho<2>, grd, blk, 0, st, a);
  else hipLaunchKernelGGL(k_rho<3>, grd, blk, 0, st, a);
  PXR_LAUNCHED();
  return finish_checked("pxr_depose_rho_n");
}

static int mk_arr3(Arr3 &a, const double *p, const int64_t *lo, const int64_t *hi) {
  if (!p || !lo || !hi) return set_error(PXRHIP_ERR_ARG, "null array or bounds");
  for (int d = 0; d < 3; d++)
    if (hi[d] < lo[d]) return set_error(PXRHIP_ERR_ARG, "array bound hi < lo");
  a.p = const_cast<double *>(p);
  a.lo0 = (int)lo[0];
  a.lo1 = (int)lo[1];
  a.lo2 = (int)lo[2];
  a.sy = hi[0] - lo[0] + 1;
  a.sz = a.sy * (hi[1] - lo[1] + 1);
  return PXRHIP_OK;
}

// every node the loop box touches, with its stencil neighbours, must lie in
// the arrays: Yee E reads B at -1, B reads E at +1
static int check_box(const Box &b, const int64_t *alo, const int64_t *ahi, int dm, int dp) {
  for (int d = 0; d < 3; d++)
    if (b.lo[d] + dm < alo[d] || b.hi[d] + dp > ahi[d])
      return set_error(PXRHIP_ERR_ARG, "FDTD loop box leaves an array's bounds");
  return PXRHIP_OK;
}

static void mk_box(Box &b, const int64_t *lo, const int64_t *hi) {
  for (int d = 0; d < 3; d++) {
    b.lo[d] = (int)lo[d];
    b.hi[d] = (int)hi[d];
  }
}

static void union_box(YeeArgs &a) {
  for (int d = 0; d < 3; d++) {
    a.un.lo[d] = min(a.bx[0].lo[d], min(a.bx[1].lo[d], a.bx[2].lo[d]));
    a.un.hi[d] = max(a.bx[0].hi[d], max(a.bx[1].hi[d], a.bx[2].hi[d]));
  }
}

int pxrhip_push_em3d_evec(const int64_t *xlo, const int64_t *xhi, const int64_t *ylo,
                          const int64_t *yhi, const int64_t *zlo, const int64_t *zhi, double *ex,
                          const int64_t *exlo, const int64_t *exhi, double *ey, const int64_t *eylo,
                          const int64_t *eyhi, double *ez, const int64_t *ezlo, const int64_t *ezhi,
                          const double *bx, const int64_t *bxlo, const int64_t *bxhi,
                          const double *by, const int64_t *bylo, const int64_t *byhi,
                          const double *bz, const int64_t *bzlo, const int64_t *bzhi,
                          const double *jx, const int64_t *jxlo, const int64_t *jxhi,
                          const double *jy, const int64_t *jylo, const int64_t *jyhi,
                          const double *jz, const int64_t *jzlo, const int64_t *jzhi, double mudt,
                          double dtsdx, double dtsdy, double dtsdz) {
  PXR_INIT();
  if (!xlo || !xhi || !ylo || !yhi || !zlo || !zhi) return set_error(PXRHIP_ERR_ARG, "null loop bounds");
  YeeArgs a;
  int r;
  const double *P[9] = {ex, ey, ez, bx, by, bz, jx, jy, jz};
  const int64_t *L[9] = {exlo, eylo, ezlo, bxlo, bylo, bzlo, jxlo, jylo, jzlo};
  const int64_t *H[9] = {exhi, eyhi, ezhi, bxhi, byhi, bzhi, jxhi, jyhi, jzhi};
  for (int c = 0; c < 9; c++)
    if ((r = mk_arr3(a.f[c], P[c], L[c], H[c]))) return r;
  mk_box(a.bx[0], xlo, xhi);
  mk_box(a.bx[1], ylo, yhi);
  mk_box(a.bx[2], zlo, zhi);
  for (int c = 0; c < 3; c++) {
    for (int d = 0; d < 3; d++)
      if (a.bx[c].hi[d] < a.bx[c].lo[d]) return PXRHIP_OK;  // empty loop
    // the component itself and its J at the box; the curl reaches -1
    if ((r = check_box(a.bx[c], L[c], H[c], 0, 0))) return r;
    if ((r = check_box(a.bx[c], L[6 + c], H[6 + c], 0, 0))) return r;
    for (int b = 3; b < 6; b++)
      if (b - 3 != c && (r = check_box(a.bx[c], L[b], H[b], -1, 0))) return r;
  }
  union_box(a);
  a.mudt = mudt;
  a.cx = dtsdx;
  a.cy = dtsdy;
  a.cz = dtsdz;
  i64 tot = (i64)(a.un.hi[0] - a.un.lo[0] + 1) * (a.un.hi[1] - a.un.lo[1] + 1) *
            (a.un.hi[2] - a.un.lo[2] + 1);
  hipLaunchKernelGGL(k_yee_e, dim3(grid1d(tot, 256)), dim3(256), 0, rt().stream, a);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_push_em3d_bvec(const int64_t *xlo, const int64_t *xhi, const int64_t *ylo,
                          const int64_t *yhi, const int64_t *zlo, const int64_t *zhi,
                          const double *ex, const int64_t *exlo, const int64_t *exhi,
                          const double *ey, const int64_t *eylo, const int64_t *eyhi,
                          const double *ez, const int64_t *ezlo, const int64_t *ezhi, double *bx,
                          const int64_t *bxlo, const int64_t *bxhi, double *by, const int64_t *bylo,
                          const int64_t *byhi, double *bz, const int64_t *bzlo, const int64_t *bzhi,
                          double dtsdx, double dtsdy, double dtsdz) {
  PXR_INIT();
  if (!xlo || !xhi || !ylo || !yhi || !zlo || !zhi) return set_error(PXRHIP_ERR_ARG, "null loop bounds");
  YeeArgs a;
  int r;
  const double *P[6] = {ex, ey, ez, bx, by, bz};
  const int64_t *L[6] = {exlo, eylo, ezlo, bxlo, bylo, bzlo};
  const int64_t *H[6] = {exhi, eyhi, ezhi, bxhi, byhi, bzhi};
  for (int c = 0; c < 6; c++)
    if ((r = mk_arr3(a.f[c], P[c], L[c], H[c]))) return r;
  for (int c = 6; c < 9; c++) a.f[c] = a.f[0];
  mk_box(a.bx[0], xlo, xhi);
  mk_box(a.bx[1], ylo, yhi);
  mk_box(a.bx[2], zlo, zhi);
  for (int c = 0; c < 3; c++) {
    for (int d = 0; d < 3; d++)
      if (a.bx[c].hi[d] < a.bx[c].lo[d]) return PXRHIP_OK;
    if ((r = check_box(a.bx[c], L[3 + c], H[3 + c], 0, 0))) return r;
    for (int e = 0; e < 3; e++)
      if (e != c && (r = check_box(a.bx[c], L[e], H[e], 0, 1))) return r;
  }
  union_box(a);
  a.mudt = 0.0;
  a.cx = dtsdx;
  a.cy = dtsdy;
  a.cz = dtsdz;
  i64 tot = (i64)(a.un.hi[0] - a.un.lo[0] + 1) * (a.un.hi[1] - a.un.lo[1] + 1) *
            (a.un.hi[2] - a.un.lo[2] + 1);
  hipLaunchKernelGGL(k_yee_b, dim3(grid1d(tot, 256)), dim3(256), 0, rt().stream, a);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

static int mk_norder(NorderArgs &a, const double *cx, const double *cy, const double *cz,
                     int64_t nx, int64_t ny, int64_t nz, int64_t nox, int64_t noy, int64_t noz,
                     int64_t gx, int64_t gy, int64_t gz, int64_t nxs, int64_t nys, int64_t nzs,
                     int64_t l_nodal) {
  int r = check_grid(nx, ny, nz, gx, gy, gz);
  if (r) return r;
  const int64_t no[3] = {nox, noy, noz};
  const double *cc[3] = {cx, cy, cz};
  for (int d = 0; d < 3; d++) {
    if (no[d] < 2 || no[d] % 2 || no[d] / 2 > PXR_MAX_HALF_ORDER)
      return set_error(PXRHIP_ERR_ARG, "stencil order must be even, 2..%d", 2 * PXR_MAX_HALF_ORDER);
    if (!cc[d]) return set_error(PXRHIP_ERR_ARG, "null stencil coefficients");
    a.co.nh[d] = (int)(no[d] / 2);
    for (int i = 0; i < a.co.nh[d]; i++) a.co.c[d][i] = cc[d][i];
  }
  // the loop box -s..n+s with a stencil reach of at most the guard width
  if (nxs < 0 || nys < 0 || nzs < 0 || nxs > gx || nys > gy || nzs > gz)
    return set_error(PXRHIP_ERR_ARG, "nxs..nzs must lie in 0..guards");
  a.g = mkgeom(nx, ny, nz, gx, gy, gz);
  a.nxs = (int)nxs;
  a.nys = (int)nys;
  a.nzs = (int)nzs;
  a.ist = l_nodal ? 0 : 1;
  return PXRHIP_OK;
}

int pxrhip_push_em3d_evec_norder(double *ex, double *ey, double *ez, const double *bx,
                                 const double *by, const double *bz, const double *jx,
                                 const double *jy, const double *jz, double mudt,
                                 const double *dtsdx, const double *dtsdy, const double *dtsdz,
                                 int64_t nx, int64_t ny, int64_t nz, int64_t norderx,
                                 int64_t nordery, int64_t norderz, int64_t nxguard,
                                 int64_t nyguard, int64_t nzguard, int64_t nxs, int64_t nys,
                                 int64_t nzs, int64_t l_nodalgrid) {
  PXR_INIT();
  NorderArgs a;
  int r = mk_norder(a, dtsdx, dtsdy, dtsdz, nx, ny, nz, norderx, nordery, norderz, nxguard,
                    nyguard, nzguard, nxs, nys, nzs, l_nodalgrid);
  if (r) return r;
  double *f[9] = {ex, ey, ez, const_cast<double *>(bx), const_cast<double *>(by),
                  const_cast<double *>(bz), const_cast<double *>(jx), const_cast<double *>(jy),
                  const_cast<double *>(jz)};
  for (int c = 0; c < 9; c++) {
    if (!f[c]) return set_error(PXRHIP_ERR_ARG, "null array");
    a.f[c] = f[c];
  }
  a.mudt = mudt;
  i64 tot = (i64)(nx + 2 * nxs + 1) * (ny + 2 * nys + 1) * (nz + 2 * nzs + 1);
  launch_norder(a, true, tot);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_push_em3d_bvec_norder(const double *ex, const double *ey, const double *ez, double *bx,
                                 double *by, double *bz, const double *dtsdx, const double *dtsdy,
                                 const double *dtsdz, int64_t nx, int64_t ny, int64_t nz,
                                 int64_t norderx, int64_t nordery, int64_t norderz,
                                 int64_t nxguard, int64_t nyguard, int64_t nzguard, int64_t nxs,
                                 int64_t nys, int64_t nzs, int64_t l_nodalgrid) {
  PXR_INIT();
  NorderArgs a;
  int r = mk_norder(a, dtsdx, dtsdy, dtsdz, nx, ny, nz, norderx, nordery, norderz, nxguard,
                    nyguard, nzguard, nxs, nys, nzs, l_nodalgrid);
  if (r) return r;
  double *f[6] = {const_cast<double *>(ex), const_cast<double *>(ey), const_cast<double *>(ez),
                  bx, by, bz};
  for (int c = 0; c < 6; c++) {
    if (!f[c]) return set_error(PXRHIP_ERR_ARG, "null array");
    a.f[c] = f[c];
  }
  for (int c = 6; c < 9; c++) a.f[c] = nullptr;
  a.mudt = 0.0;
  i64 tot = (i64)(nx + 2 * nxs + 1) * (ny + 2 * nys + 1) * (nz + 2 * nzs + 1);
  launch_norder(a, false, tot);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_field_bc_periodic(double *f, int64_t nx, int64_t ny, int64_t nz, int64_t nxg,
                             int64_t nyg, int64_t nzg) {
  PXR_INIT();
  int r = check_grid(nx, ny, nz, nxg, nyg, nzg);
  if (r) return r;
  if (!f) return set_error(PXRHIP_ERR_ARG, "null array");
  if (nx <= 2 * nxg || ny <= 2 * nyg || nz <= 2 * nzg)
    return set_error(PXRHIP_ERR_UNSUPPORTED, "periodic guards need n > 2*guards on every axis");
  GridGeom g = mkgeom(nx, ny, nz, nxg, nyg, nzg);
  const int64_t ext[3] = {nx + 1 + 2 * nxg, ny + 1 + 2 * nyg, nz + 1 + 2 * nzg};
  const int64_t gg[3] = {nxg, nyg, nzg};
  for (int ax = 0; ax < 3; ax++) {
    int64_t tot = 1;
    for (int d = 0; d < 3; d++) tot *= (d == ax) ? 2 * gg[d] + 1 : ext[d];
    hipLaunchKernelGGL(k_field_bc_axis, dim3(grid1d(tot, 256)), dim3(256), 0, rt().stream, f, g, ax);
    PXR_LAUNCHED();
  }
  return PXRHIP_OK;
}

int pxrhip_current_bc_periodic(double *a, int64_t nx, int64_t ny, int64_t nz, int64_t nxg,
                               int64_t nyg, int64_t nzg) {
  PXR_INIT();
  int r = check_grid(nx, ny, nz, nxg, nyg, nzg);
  if (r) return r;
  if (!a) return set_error(PXRHIP_ERR_ARG, "null array");
  if (nx <= 2 * nxg || ny <= 2 * nyg || nz <= 2 * nzg)
    return set_error(PXRHIP_ERR_UNSUPPORTED, "periodic guards need n > 2*guards on every axis");
  GridGeom g = mkgeom(nx, ny, nz, nxg, nyg, nzg);
  const int64_t ext[3] = {nx + 1 + 2 * nxg, ny + 1 + 2 * nyg, nz + 1 + 2 * nzg};
  const int64_t gg[3] = {nxg, nyg, nzg};
  for (int ax = 0; ax < 3; ax++) {
    int64_t tot = 1;
    for (int d = 0; d < 3; d++) tot *= (d == ax) ? 2 * gg[d] + 1 : ext[d];
    hipLaunchKernelGGL(k_current_bc_axis, dim3(grid1d(tot, 256)), dim3(256), 0, rt().stream, a, g, ax);
    PXR_LAUNCHED();
  }
  return PXRHIP_OK;
}

}  // extern "C"

// Arguments of pxr_particle_bin_sorting checked and turned into the sorting
// grid (sorting.F90:274-278).  Returns PXRHIP_OK with *empty set when there
// is nothing to sort.
static int bin_sort_geom(int64_t np2, double *const *arr, int64_t wpid, double xmin2,
                         double ymin2, double zmin2, double xmax2, double ymax2, double zmax2,
                         double dxf, double dyf, double dzf, SortGeom &s, bool &empty) {
  empty = true;
  if (np2 < 0) return set_error(PXRHIP_ERR_ARG, "np2 < 0");
  if (wpid != 1) return set_error(PXRHIP_ERR_ARG, "only the first pid column (wpid=1) is permuted");
  if (np2 == 0) return PXRHIP_OK;
  if (np2 > INT32_MAX) return set_error(PXRHIP_ERR_ARG, "np2 too large for one sort");
  for (int c = 0; c < 8; c++)
    if (!arr[c]) return set_error(PXRHIP_ERR_ARG, "null array");
  s.xmin = xmin2; s.ymin = ymin2; s.zmin = zmin2;
  s.dxi = 1.0 / dxf; s.dyi = 1.0 / dyf; s.dzi = 1.0 / dzf;  // sorting.F90:274-276
  s.nx3 = (i64)ceil((xmax2 - xmin2) * s.dxi);
  s.ny3 = (i64)ceil((ymax2 - ymin2) * s.dyi);
  const i64 nz3 = (i64)ceil((zmax2 - zmin2) * s.dzi);
  if (s.nx3 <= 0 || s.ny3 <= 0 || nz3 <= 0) return set_error(PXRHIP_ERR_ARG, "empty sorting grid");
  const double nb = (double)s.nx3 * (double)s.ny3 * (double)nz3;
  if (nb > 1.0e8) return set_error(PXRHIP_ERR_ARG, "sorting grid has more than 1e8 bins");
  s.nbins = s.nx3 * s.ny3 * nz3;
  empty = false;
  return PXRHIP_OK;
}

// The sort itself, on the library stream, with the caller's buffers: key
// and dest hold np2 ints, tmp np2 doubles, hist s.nbins ints, bad one
// counter that is zero on entry and left zero on return.  Keys are checked
// before anything moves (the host reads the counter, so the call
// synchronises the stream once).
static int bin_sort_run(int64_t np2, double *const *arr, const SortGeom &s, int *key, int *hist,
                        int *dest, double *tmp, unsigned long long *bad_ctr) {
  hipStream_t st = rt().stream;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid1d(s.nbins, 256)), dim3(256), 0, st, hist, s.nbins, 0);
  hipLaunchKernelGGL(k_sort_keys, dim3(grid1d(np2, 256)), dim3(256), 0, st, np2, arr[0], arr[1],
                     arr[2], s, key, hist, bad_ctr);
  PXR_LAUNCHED();
  unsigned long long bad = 0;
  PXR_HIP(hipMemcpyAsync(&bad, bad_ctr, sizeof(bad), hipMemcpyDeviceToHost, st));
  PXR_HIP(hipStreamSynchronize(st));
  if (bad) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, bad_ctr, 1, 0ull);
    PXR_LAUNCHED();
    return set_error(PXRHIP_ERR_BOUNDS,
                     "particle_bin_sorting: %llu particle(s) have a cell key outside [1,%lld]; "
                     "nothing was moved", bad, (long long)s.nbins);
  }
  hipLaunchKernelGGL(k_scan_1wg, dim3(1), dim3(1024), 0, st, hist, s.nbins);
  hipLaunchKernelGGL(k_sort_scatter_1wg, dim3(1), dim3(1024), 0, st, np2, key, hist, dest);
  PXR_LAUNCHED();
  for (int c = 0; c < 8; c++) {
    hipLaunchKernelGGL(k_permute, dim3(grid1d(np2, 256)), dim3(256), 0, st, np2, arr[c], tmp, dest);
    PXR_HIP(hipMemcpyAsync(arr[c], tmp, sizeof(double) * np2, hipMemcpyDeviceToDevice, st));
  }
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

// Fixed device scratch of pxrhip_particle_bin_sorting_noalloc: key, dest,
// tmp and hist of one sort of at most SORT_NA_NP particles over at most
// SORT_NA_NBINS bins, and the bad-key counter.  Each GPU holds its own copy
// until the process ends; calls on one GPU share it.
constexpr i64 SORT_NA_NP = 1 << 20;
constexpr i64 SORT_NA_NBINS = 1 << 20;
constexpr size_t SORT_NA_OFF_DEST = sizeof(int) * SORT_NA_NP;
constexpr size_t SORT_NA_OFF_TMP = SORT_NA_OFF_DEST + sizeof(int) * SORT_NA_NP;
constexpr size_t SORT_NA_OFF_HIST = SORT_NA_OFF_TMP + sizeof(double) * SORT_NA_NP;
constexpr size_t SORT_NA_OFF_BAD = SORT_NA_OFF_HIST + sizeof(int) * SORT_NA_NBINS;
constexpr size_t SORT_NA_BYTES = SORT_NA_OFF_BAD + 256;
static __device__ __align__(256) unsigned char sort_scratch[SORT_NA_BYTES];

// Device address of sort_scratch on each device (the symbol's name in host
// code is not a device address), looked up on the first call on a device.
constexpr int SORT_NA_MAXDEV = 64;
static unsigned char *g_sort_scratch[SORT_NA_MAXDEV] = {};

static int sort_scratch_ptr(unsigned char **out) {
  const int dev = rt().device;
  if (dev < 0 || dev >= SORT_NA_MAXDEV)
    return set_error(PXRHIP_ERR_ARG, "device %d out of range for the fixed sort scratch", dev);
  if (!g_sort_scratch[dev]) {
    void *p = nullptr;
    PXR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(sort_scratch)));
    // the counter starts at zero: a fill kernel, not a memset, so that
    // work recorded into a graph clears it the same way
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, rt().stream,
                       (unsigned long long *)((unsigned char *)p + SORT_NA_OFF_BAD), 1, 0ull);
    PXR_LAUNCHED();
    g_sort_scratch[dev] = (unsigned char *)p;
  }
  *out = g_sort_scratch[dev];
  return PXRHIP_OK;
}

extern "C" {

int pxrhip_particle_bin_sorting(int64_t np2, double *xp, double *yp, double *zp, double *ux,
                                double *uy, double *uz, double *gam, double *pid, int64_t wpid,
                                double xmin2, double ymin2, double zmin2, double xmax2,
                                double ymax2, double zmax2, double dxf, double dyf, double dzf) {
  PXR_INIT();
  double *arr[8] = {xp, yp, zp, ux, uy, uz, gam, pid};
  SortGeom s;
  bool empty;
  int r = bin_sort_geom(np2, arr, wpid, xmin2, ymin2, zmin2, xmax2, ymax2, zmax2, dxf, dyf, dzf, s,
                        empty);
  if (r || empty) return r;
  int *key = (int *)scratch(0, sizeof(int) * np2);
  int *hist = (int *)scratch(1, sizeof(int) * s.nbins);
  int *dest = (int *)scratch(2, sizeof(int) * np2);
  double *tmp = (double *)scratch(3, sizeof(double) * np2);
  if (!key || !hist || !dest || !tmp) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
  return bin_sort_run(np2, arr, s, key, hist, dest, tmp, rt().d_bad);
}

int pxrhip_particle_bin_sorting_noalloc(int64_t np2, double *xp, double *yp, double *zp,
                                        double *ux, double *uy, double *uz, double *gam,
                                        double *pid, int64_t wpid, double xmin2, double ymin2,
                                        double zmin2, double xmax2, double ymax2, double zmax2,
                                        double dxf, double dyf, double dzf) {
  // no implicit set-up here: pxrhip_init allocates, this entry point may not
  if (rt().device < 0)
    return set_error(PXRHIP_ERR_ARG, "call pxrhip_init before pxrhip_particle_bin_sorting_noalloc");
  double *arr[8] = {xp, yp, zp, ux, uy, uz, gam, pid};
  SortGeom s;
  bool empty;
  int r = bin_sort_geom(np2, arr, wpid, xmin2, ymin2, zmin2, xmax2, ymax2, zmax2, dxf, dyf, dzf, s,
                        empty);
  if (r || empty) return r;
  if (np2 > SORT_NA_NP)
    return set_error(PXRHIP_ERR_ARG, "np2 = %lld exceeds the %lld particles of the fixed sort scratch",
                     (long long)np2, (long long)SORT_NA_NP);
  if (s.nbins > SORT_NA_NBINS)
    return set_error(PXRHIP_ERR_ARG, "the sorting grid has %lld bins, more than the %lld of the "
                     "fixed sort scratch", (long long)s.nbins, (long long)SORT_NA_NBINS);
  unsigned char *base;
  if ((r = sort_scratch_ptr(&base))) return r;
  return bin_sort_run(np2, arr, s, (int *)base, (int *)(base + SORT_NA_OFF_HIST),
                      (int *)(base + SORT_NA_OFF_DEST), (double *)(base + SORT_NA_OFF_TMP),
                      (unsigned long long *)(base + SORT_NA_OFF_BAD));
}

}  // extern "C"
