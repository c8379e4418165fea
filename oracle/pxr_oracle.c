/*
 * pxr_oracle.c -- CPU FP64 restatement of PICSAR's per-step PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library
 * is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * Every routine restates the algorithm of the reference routine cited in
 * its comment (paths relative to the PICSAR source tree).  Arithmetic is
 * written in the reference's evaluation order and the file is compiled with
 * -ffp-contract=off, so that integer cell indices computed here are the
 * reference's own.
 *
 * Conventions (from the reference's Fortran declarations):
 *   - particle arrays are contiguous float64 of length np;
 *   - grid arrays are Fortran-ordered float64 with bounds
 *     (-ngx:nx+ngx, -ngy:ny+ngy, -ngz:nz+ngz); element (i,j,k) lives at
 *     (i+ngx) + NX*((j+ngy) + NY*(k+ngz)), NX = nx+1+2*ngx, NY = ny+1+2*ngy;
 *   - integer scalars are int64, logicals are passed as int64 0/1.
 *
 * Each function is pinned by the tests in tests/test_oracle_golden.py
 * (vectors produced by the reference's own compiled kernels) and
 * tests/test_oracle_math.py (closed forms and invariants).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

#define ONESIXTH (1.0 / 6.0)
#define TWOTHIRD (2.0 / 3.0)
#define ONETHIRD (1.0 / 3.0)

typedef struct {
  i64 nx, ny, nz, gx, gy, gz, NX, NY;
} grid_t;

static grid_t mkgrid(i64 nx, i64 ny, i64 nz, i64 gx, i64 gy, i64 gz) {
  grid_t g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.gx = gx; g.gy = gy; g.gz = gz;
  g.NX = nx + 1 + 2 * gx;
  g.NY = ny + 1 + 2 * gy;
  return g;
}

static inline i64 gidx(const grid_t *g, i64 i, i64 j, i64 k) {
  return (i + g->gx) + g->NX * ((j + g->gy) + g->NY * (k + g->gz));
}

static inline int ginside(const grid_t *g, i64 i, i64 j, i64 k) {
  return i >= -g->gx && i <= g->nx + g->gx && j >= -g->gy &&
         j <= g->ny + g->gy && k >= -g->gz && k <= g->nz + g->gz;
}

/* Fortran NINT: round half away from zero. */
static inline i64 f_nint(double x) { return (i64)round(x); }
static inline i64 f_floor(double x) { return (i64)floor(x); }

/* B-spline weights as written in the reference kernels.
 *   linear : field_gathering_o1_3d.F90:145-150
 *   quad   : field_gathering_o2_3d.F90:147-158
 *   cubic  : field_gathering_o3_3d.F90:175-201                         */
static inline void w_lin(double f, double *w) {
  w[0] = 1.0 - f;
  w[1] = f;
}
static inline void w_quad(double f, double *w) {
  double a = 0.5 - f, b = 0.5 + f;
  w[0] = 0.5 * (a * a);
  w[1] = 0.75 - f * f;
  w[2] = 0.5 * (b * b);
}
static inline void w_cubic(double f, double *w) {
  double o = 1.0 - f, f2 = f * f, o2 = o * o;
  w[0] = ONESIXTH * o2 * o;
  w[1] = TWOTHIRD - f2 * (1.0 - f * 0.5);
  w[2] = TWOTHIRD - o2 * (1.0 - o * 0.5);
  w[3] = ONESIXTH * f2 * f;
}

/* Node-centred shape of order n at normalised coordinate x (gather).
 * Order 1/3 use floor, order 2 uses nint (field_gathering_o2_3d.F90:141).
 * Returns the first node; w[0..n] are the weights.                     */
static inline i64 shape_nodal(int n, double x, double *w) {
  i64 j;
  switch (n) {
  case 1: j = f_floor(x); w_lin(x - (double)j, w); return j;
  case 2: j = f_nint(x); w_quad(x - (double)j, w); return j - 1;
  default: j = f_floor(x); w_cubic(x - (double)j, w); return j - 1;
  }
}

/* Staggered shape of order m (0..3) used by the gather for the component
 * on the half-integer grid, shift s = 0.5 (Yee) or 0 (nodal grid).
 * Index rules: m even -> floor(x+0.5-s); m odd -> floor(x-s)
 * (field_gathering_o1_3d.F90:139-141, o2:137-143, o3:150-156/301-307). */
static inline i64 shape_stag(int m, double x, double s, double *w) {
  i64 j0;
  switch (m) {
  case 0: j0 = f_floor((x + 0.5) - s); w[0] = 1.0; return j0;
  case 1: j0 = f_floor(x - s); w_lin((x - s) - (double)j0, w); return j0;
  case 2: j0 = f_floor((x + 0.5) - s); w_quad((x - s) - (double)j0, w); return j0 - 1;
  default: j0 = f_floor(x - s); w_cubic((x - s) - (double)j0, w); return j0 - 1;
  }
}

/* ------------------------------------------------------------------------
 * Field gathering, energy conserving, orders 1..3.
 * Restates gete3d/getb3d_energy_conserving_scalar_{1,2,3} (field_gathering_
 * o1_3d.F90:74/328, o2_3d.F90:70/366, o3_3d.F90:83/406) as dispatched from
 * geteb3d_energy_conserving (field_gathering_manager_3d.F90:295).  Results
 * are ADDED into ex..bz (the reference accumulates into caller-zeroed
 * arrays).  Returns the number of particles whose stencil leaves the array
 * (those particles are skipped; the reference reads out of bounds).      */
i64 orc_geteb3d_energy_conserving(
    i64 np, const double *xp, const double *yp, const double *zp, double *ex,
    double *ey, double *ez, double *bx, double *by, double *bz, double xmin,
    double ymin, double zmin, double dx, double dy, double dz, i64 nx, i64 ny,
    i64 nz, i64 ngx, i64 ngy, i64 ngz, i64 nox, i64 noy, i64 noz,
    const double *exg, const double *eyg, const double *ezg, const double *bxg,
    const double *byg, const double *bzg, i64 l_lower_order_in_v, i64 l_nodal) {
  if (nox != noy || nox != noz || nox < 1 || nox > 3) return -1;
  const int n = (int)nox;
  const int m = l_lower_order_in_v ? n - 1 : n;
  const double s = l_nodal ? 0.0 : 0.5;
  const double dxi = 1.0 / dx, dyi = 1.0 / dy, dzi = 1.0 / dz;
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  i64 bad = 0;
  for (i64 ip = 0; ip < np; ip++) {
    double x = (xp[ip] - xmin) * dxi;
    double y = (yp[ip] - ymin) * dyi;
    double z = (zp[ip] - zmin) * dzi;
    double sx[4], sy[4], sz[4], sx0[4], sy0[4], sz0[4];
    i64 j = shape_nodal(n, x, sx), k = shape_nodal(n, y, sy),
        l = shape_nodal(n, z, sz);
    i64 j0 = shape_stag(m, x, s, sx0), k0 = shape_stag(m, y, s, sy0),
        l0 = shape_stag(m, z, s, sz0);
    i64 lo_i = j < j0 ? j : j0, lo_j = k < k0 ? k : k0, lo_k = l < l0 ? l : l0;
    i64 hi_i = (j + n > j0 + m) ? j + n : j0 + m;
    i64 hi_j = (k + n > k0 + m) ? k + n : k0 + m;
    i64 hi_k = (l + n > l0 + m) ? l + n : l0 + m;
    if (!ginside(&g, lo_i, lo_j, lo_k) || !ginside(&g, hi_i, hi_j, hi_k)) {
      bad++;
      continue;
    }
    /* Ex: x staggered, y/z nodal */
    for (int c = 0; c <= n; c++)
      for (int b = 0; b <= n; b++)
        for (int a = 0; a <= m; a++)
          ex[ip] = ex[ip] + sx0[a] * sy[b] * sz[c] * exg[gidx(&g, j0 + a, k + b, l + c)];
    /* Ey: y staggered */
    for (int c = 0; c <= n; c++)
      for (int b = 0; b <= m; b++)
        for (int a = 0; a <= n; a++)
          ey[ip] = ey[ip] + sx[a] * sy0[b] * sz[c] * eyg[gidx(&g, j + a, k0 + b, l + c)];
    /* Ez: z staggered */
    for (int c = 0; c <= m; c++)
      for (int b = 0; b <= n; b++)
        for (int a = 0; a <= n; a++)
          ez[ip] = ez[ip] + sx[a] * sy[b] * sz0[c] * ezg[gidx(&g, j + a, k + b, l0 + c)];
    /* Bx: y/z staggered */
    for (int c = 0; c <= m; c++)
      for (int b = 0; b <= m; b++)
        for (int a = 0; a <= n; a++)
          bx[ip] = bx[ip] + sx[a] * sy0[b] * sz0[c] * bxg[gidx(&g, j + a, k0 + b, l0 + c)];
    /* By: x/z staggered */
    for (int c = 0; c <= m; c++)
      for (int b = 0; b <= n; b++)
        for (int a = 0; a <= m; a++)
          by[ip] = by[ip] + sx0[a] * sy[b] * sz0[c] * byg[gidx(&g, j0 + a, k + b, l0 + c)];
    /* Bz: x/y staggered */
    for (int c = 0; c <= n; c++)
      for (int b = 0; b <= m; b++)
        for (int a = 0; a <= m; a++)
          bz[ip] = bz[ip] + sx0[a] * sy0[b] * sz[c] * bzg[gidx(&g, j0 + a, k0 + b, l + c)];
  }
  return bad;
}

/* ------------------------------------------------------------------------
 * Boris momentum push, boris_3d.F90:51-141 (pxr_boris_push_u_3d).
 * u is gamma*v in velocity units; clight = 1 (normalised units).         */
void orc_boris_push_u_3d(i64 np, double *uxp, double *uyp, double *uzp,
                         double *gaminv, const double *ex, const double *ey,
                         const double *ez, const double *bx, const double *by,
                         const double *bz, double q, double m, double dt,
                         double clight) {
  const double cst = q * dt * 0.5 / m;
  const double clghtisq = 1.0 / (clight * clight);
  for (i64 ip = 0; ip < np; ip++) {
    double ux = uxp[ip] + ex[ip] * cst;
    double uy = uyp[ip] + ey[ip] * cst;
    double uz = uzp[ip] + ez[ip] * cst;
    double usq = (ux * ux + uy * uy + uz * uz) * clghtisq;
    double gi = 1.0 / sqrt(1.0 + usq);
    double tx = gi * bx[ip] * cst, ty = gi * by[ip] * cst, tz = gi * bz[ip] * cst;
    double tsqi = 2.0 / (1.0 + tx * tx + ty * ty + tz * tz);
    double sx = tx * tsqi, sy = ty * tsqi, sz = tz * tsqi;
    double uxp_ = ux + uy * tz - uz * ty;
    double uyp_ = uy + uz * tx - ux * tz;
    double uzp_ = uz + ux * ty - uy * tx;
    ux = ux + uyp_ * sz - uzp_ * sy;
    uy = uy + uzp_ * sx - uxp_ * sz;
    uz = uz + uxp_ * sy - uyp_ * sx;
    ux = ux + ex[ip] * cst;
    uy = uy + ey[ip] * cst;
    uz = uz + ez[ip] * cst;
    usq = (ux * ux + uy * uy + uz * uz) * clghtisq;
    gaminv[ip] = 1.0 / sqrt(1.0 + usq);
    uxp[ip] = ux; uyp[ip] = uy; uzp[ip] = uz;
  }
}

/* pxr_set_gamma, boris_3d.F90:1059-1107. */
void orc_set_gamma(i64 np, const double *uxp, const double *uyp,
                   const double *uzp, double *gaminv, double clight) {
  const double clghtisq = 1.0 / (clight * clight);
  for (i64 ip = 0; ip < np; ip++) {
    double usq = (uxp[ip] * uxp[ip] + uyp[ip] * uyp[ip] + uzp[ip] * uzp[ip]) * clghtisq;
    gaminv[ip] = 1.0 / sqrt(1.0 + usq);
  }
}

/* pxr_pushxyz, boris_3d.F90:849-893: x += u*gaminv*dt. */
void orc_pushxyz(i64 np, double *xp, double *yp, double *zp, const double *uxp,
                 const double *uyp, const double *uzp, const double *gaminv,
                 double dt) {
  for (i64 ip = 0; ip < np; ip++) {
    xp[ip] = xp[ip] + uxp[ip] * gaminv[ip] * dt;
    yp[ip] = yp[ip] + uyp[ip] * gaminv[ip] * dt;
    zp[ip] = zp[ip] + uzp[ip] * gaminv[ip] * dt;
  }
}

/* Vay pusher, vay_3d.F90:56-241 (pxr_ebcancelpush3d); which = 0 full push,
 * 1 first half (E + v x B with old gamma), 2 second half.               */
void orc_ebcancelpush3d(i64 np, double *uxp, double *uyp, double *uzp,
                        double *gi, const double *ex, const double *ey,
                        const double *ez, const double *bx, const double *by,
                        const double *bz, double q, double m, double dt,
                        i64 which, double clight) {
  const double invclight = 1.0 / clight;
  const double invclightsq = 1.0 / (clight * clight);
  if (which == 0 || which == 2) {
    const double cst = which == 0 ? q * dt / m : 0.5 * q * dt / m;
    const double bconst = which == 0 ? 0.5 * cst : cst;
    for (i64 ip = 0; ip < np; ip++) {
      double taux = bconst * bx[ip], tauy = bconst * by[ip], tauz = bconst * bz[ip];
      double tausq = taux * taux + tauy * tauy + tauz * tauz;
      double uxpr, uypr, uzpr;
      if (which == 0) {
        uxpr = uxp[ip] + cst * ex[ip] + (uyp[ip] * tauz - uzp[ip] * tauy) * gi[ip];
        uypr = uyp[ip] + cst * ey[ip] + (uzp[ip] * taux - uxp[ip] * tauz) * gi[ip];
        uzpr = uzp[ip] + cst * ez[ip] + (uxp[ip] * tauy - uyp[ip] * taux) * gi[ip];
      } else {
        uxpr = uxp[ip] + cst * ex[ip];
        uypr = uyp[ip] + cst * ey[ip];
        uzpr = uzp[ip] + cst * ez[ip];
      }
      double gprsq = (1.0 + (uxpr * uxpr + uypr * uypr + uzpr * uzpr) * invclightsq);
      double ust = (uxpr * taux + uypr * tauy + uzpr * tauz) * invclight;
      double sigma = gprsq - tausq;
      double gisq = 2.0 / (sigma + sqrt(sigma * sigma + 4.0 * (tausq + ust * ust)));
      double g = sqrt(gisq);
      gi[ip] = g;
      double bg = bconst * g;
      double tx = bg * bx[ip], ty = bg * by[ip], tz = bg * bz[ip];
      double sfac = 1.0 / (1.0 + tausq * gisq);
      double tu = tx * uxpr + ty * uypr + tz * uzpr;
      uxp[ip] = sfac * (uxpr + tx * tu + uypr * tz - uzpr * ty);
      uyp[ip] = sfac * (uypr + ty * tu + uzpr * tx - uxpr * tz);
      uzp[ip] = sfac * (uzpr + tz * tu + uxpr * ty - uypr * tx);
    }
  } else if (which == 1) {
    const double cst = 0.5 * q * dt / m;
    for (i64 ip = 0; ip < np; ip++) {
      double vx = uxp[ip] * gi[ip], vy = uyp[ip] * gi[ip], vz = uzp[ip] * gi[ip];
      uxp[ip] = uxp[ip] + cst * (ex[ip] + vy * bz[ip] - vz * by[ip]);
      uyp[ip] = uyp[ip] + cst * (ey[ip] + vz * bx[ip] - vx * bz[ip]);
      uzp[ip] = uzp[ip] + cst * (ez[ip] + vx * by[ip] - vy * bx[ip]);
      gi[ip] = 1.0 / sqrt(1.0 + (uxp[ip] * uxp[ip] + uyp[ip] * uyp[ip] +
                                 uzp[ip] * uzp[ip]) * invclightsq);
    }
  }
}

/* ------------------------------------------------------------------------
 * Direct current deposition, depose_jxjyjz_scalar_{1,2,3}
 * (direct_current_deposition_3d.F90:99-263, 879-1166, 1992-2522).
 * Shapes are evaluated at xmid = x - dt/2 * v, full order in every
 * direction; order 2 uses nint for both node sets (:940-945).           */
static inline i64 dshape(int n, double xm, double *w) {
  i64 j;
  switch (n) {
  case 1: j = f_floor(xm); w_lin(xm - (double)j, w); return j;
  case 2: j = f_nint(xm); w_quad(xm - (double)j, w); return j - 1;
  default: j = f_floor(xm); w_cubic(xm - (double)j, w); return j - 1;
  }
}
static inline i64 dshape_s(int n, double xm, double s, double *w) {
  i64 j;
  switch (n) {
  case 1: j = f_floor(xm - s); w_lin((xm - s) - (double)j, w); return j;
  case 2: j = f_nint(xm - s); w_quad((xm - s) - (double)j, w); return j - 1;
  default: j = f_floor(xm - s); w_cubic((xm - s) - (double)j, w); return j - 1;
  }
}

i64 orc_depose_direct(double *jx, double *jy, double *jz, i64 np,
                      const double *xp, const double *yp, const double *zp,
                      const double *uxp, const double *uyp, const double *uzp,
                      const double *gaminv, const double *w, double q,
                      double xmin, double ymin, double zmin, double dt,
                      double dx, double dy, double dz, i64 nx, i64 ny, i64 nz,
                      i64 ngx, i64 ngy, i64 ngz, i64 order, i64 l_nodal) {
  if (order < 1 || order > 3) return -1;
  const int n = (int)order;
  const double s = l_nodal ? 0.0 : 0.5;
  const double dxi = 1.0 / dx, dyi = 1.0 / dy, dzi = 1.0 / dz;
  const double invvol = dxi * dyi * dzi;
  const double dts2dx = 0.5 * dt * dxi, dts2dy = 0.5 * dt * dyi, dts2dz = 0.5 * dt * dzi;
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  i64 bad = 0;
  for (i64 ip = 0; ip < np; ip++) {
    double x = (xp[ip] - xmin) * dxi, y = (yp[ip] - ymin) * dyi, z = (zp[ip] - zmin) * dzi;
    double vx = uxp[ip] * gaminv[ip], vy = uyp[ip] * gaminv[ip], vz = uzp[ip] * gaminv[ip];
    double wq = q * w[ip] * invvol;
    double wqx = wq * vx, wqy = wq * vy, wqz = wq * vz;
    double xm = x - dts2dx * vx, ym = y - dts2dy * vy, zm = z - dts2dz * vz;
    double sx[4], sy[4], sz[4], sx0[4], sy0[4], sz0[4];
    i64 j = dshape(n, xm, sx), k = dshape(n, ym, sy), l = dshape(n, zm, sz);
    i64 j0 = dshape_s(n, xm, s, sx0), k0 = dshape_s(n, ym, s, sy0), l0 = dshape_s(n, zm, s, sz0);
    i64 lo_i = j < j0 ? j : j0, lo_j = k < k0 ? k : k0, lo_k = l < l0 ? l : l0;
    i64 hi_i = (j > j0 ? j : j0) + n, hi_j = (k > k0 ? k : k0) + n, hi_k = (l > l0 ? l : l0) + n;
    if (!ginside(&g, lo_i, lo_j, lo_k) || !ginside(&g, hi_i, hi_j, hi_k)) {
      bad++;
      continue;
    }
    for (int c = 0; c <= n; c++)
      for (int b = 0; b <= n; b++)
        for (int a = 0; a <= n; a++) {
          i64 ix = gidx(&g, j0 + a, k + b, l + c);
          jx[ix] = jx[ix] + sx0[a] * sy[b] * sz[c] * wqx;
        }
    for (int c = 0; c <= n; c++)
      for (int b = 0; b <= n; b++)
        for (int a = 0; a <= n; a++) {
          i64 ix = gidx(&g, j + a, k0 + b, l + c);
          jy[ix] = jy[ix] + sx[a] * sy0[b] * sz[c] * wqy;
        }
    for (int c = 0; c <= n; c++)
      for (int b = 0; b <= n; b++)
        for (int a = 0; a <= n; a++) {
          i64 ix = gidx(&g, j + a, k + b, l0 + c);
          jz[ix] = jz[ix] + sx[a] * sy[b] * sz0[c] * wqz;
        }
  }
  return bad;
}

/* ------------------------------------------------------------------------
 * Esirkepov current deposition, depose_jxjyjz_esirkepov_{1,2,3}
 * (esirkepov_3d.F90:82-250, 1734-1990, 3263-3478).
 * Shapes at the new position (index iixp0, floor for orders 1/3, nint for
 * order 2) and at the old position x - dt*v (shifted by dix = iixp-iixp0);
 * the current is the running sum of the Esirkepov weights along each axis
 * over the box [min(0,d)+lo, max(0,d)+hi].                              */
#define EOFF 3 /* array index offset: reference index -3..3 -> 0..6 */

static inline i64 eshape(int n, double x, double *w7, int shift) {
  double ww[4];
  i64 j;
  int first;
  switch (n) {
  case 1: j = f_floor(x); w_lin(x - (double)j, ww); first = 0; break;
  case 2: j = f_nint(x); w_quad(x - (double)j, ww); first = -1; break;
  default: j = f_floor(x); w_cubic(x - (double)j, ww); first = -1; break;
  }
  for (int a = 0; a <= n; a++) w7[EOFF + first + shift + a] = ww[a];
  return j;
}

i64 orc_depose_esirkepov(double *jx, double *jy, double *jz, i64 np,
                         const double *xp, const double *yp, const double *zp,
                         const double *uxp, const double *uyp,
                         const double *uzp, const double *gaminv,
                         const double *w, double q, double xmin, double ymin,
                         double zmin, double dt, double dx, double dy,
                         double dz, i64 nx, i64 ny, i64 nz, i64 ngx, i64 ngy,
                         i64 ngz, i64 order) {
  if (order < 1 || order > 3) return -1;
  const int n = (int)order;
  const int lo = (n == 1) ? 0 : -1, hi = (n == 3) ? 2 : 1;
  const double dxi = 1.0 / dx, dyi = 1.0 / dy, dzi = 1.0 / dz;
  const double dtsdx0 = dt * dxi, dtsdy0 = dt * dyi, dtsdz0 = dt * dzi;
  const double invdtdx = 1.0 / (dt * dy * dz);
  const double invdtdy = 1.0 / (dt * dx * dz);
  const double invdtdz = 1.0 / (dt * dx * dy);
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  i64 bad = 0;
  for (i64 ip = 0; ip < np; ip++) {
    double x = (xp[ip] - xmin) * dxi, y = (yp[ip] - ymin) * dyi, z = (zp[ip] - zmin) * dzi;
    double xold, yold, zold;
    if (n == 1) { /* esirkepov_3d.F90:151-153: dtsdx0*uxp*gaminv */
      xold = x - dtsdx0 * uxp[ip] * gaminv[ip];
      yold = y - dtsdy0 * uyp[ip] * gaminv[ip];
      zold = z - dtsdz0 * uzp[ip] * gaminv[ip];
    } else {      /* esirkepov_3d.F90:3338-3343: v = u*gaminv first */
      double vx = uxp[ip] * gaminv[ip], vy = uyp[ip] * gaminv[ip], vz = uzp[ip] * gaminv[ip];
      xold = x - dtsdx0 * vx;
      yold = y - dtsdy0 * vy;
      zold = z - dtsdz0 * vz;
    }
    double wq = q * w[ip];
    double wqx = wq * invdtdx, wqy = wq * invdtdy, wqz = wq * invdtdz;
    double sx0[7] = {0}, sy0[7] = {0}, sz0[7] = {0}, sx[7] = {0}, sy[7] = {0}, sz[7] = {0};
    i64 i0 = eshape(n, x, sx0, 0), j0 = eshape(n, y, sy0, 0), k0 = eshape(n, z, sz0, 0);
    /* old-position shapes: index computed first, then placed at offset d */
    i64 i1, j1, k1;
    if (n == 2) { i1 = f_nint(xold); j1 = f_nint(yold); k1 = f_nint(zold); }
    else { i1 = f_floor(xold); j1 = f_floor(yold); k1 = f_floor(zold); }
    int dix = (int)(i1 - i0), diy = (int)(j1 - j0), diz = (int)(k1 - k0);
    if (dix < -1 || dix > 1 || diy < -1 || diy > 1 || diz < -1 || diz > 1) {
      bad++;
      continue;
    }
    eshape(n, xold, sx, dix);
    eshape(n, yold, sy, diy);
    eshape(n, zold, sz, diz);
    double dsx[7], dsy[7], dsz[7];
    for (int a = 0; a < 7; a++) {
      dsx[a] = sx[a] - sx0[a];
      dsy[a] = sy[a] - sy0[a];
      dsz[a] = sz[a] - sz0[a];
    }
    int ixmin = (dix < 0 ? dix : 0) + lo, ixmax = (dix > 0 ? dix : 0) + hi;
    int iymin = (diy < 0 ? diy : 0) + lo, iymax = (diy > 0 ? diy : 0) + hi;
    int izmin = (diz < 0 ? diz : 0) + lo, izmax = (diz > 0 ? diz : 0) + hi;
    if (!ginside(&g, i0 + ixmin, j0 + iymin, k0 + izmin) ||
        !ginside(&g, i0 + ixmax, j0 + iymax, k0 + izmax)) {
      bad++;
      continue;
    }
    double sdyj[7] = {0}, sdzk[7][7] = {{0}};
    for (int k = izmin; k <= izmax; k++) {
      for (int j = iymin; j <= iymax; j++) {
        double sdxi = 0.0;
        for (int i = ixmin; i <= ixmax; i++) {
          i64 ic = i0 + i, jc = j0 + j, kc = k0 + k;
          const int I = i + EOFF, J = j + EOFF, K = k + EOFF;
          if (i < ixmax) {
            double t = wqx * dsx[I] * ((sy0[J] + 0.5 * dsy[J]) * sz0[K] +
                                        (0.5 * sy0[J] + ONETHIRD * dsy[J]) * dsz[K]);
            sdxi = (i > ixmin) ? t + sdxi : t;
            i64 ix = gidx(&g, ic, jc, kc);
            jx[ix] = jx[ix] + sdxi;
          }
          if (j < iymax) {
            double t = wqy * dsy[J] * ((sz0[K] + 0.5 * dsz[K]) * sx0[I] +
                                        (0.5 * sz0[K] + ONETHIRD * dsz[K]) * dsx[I]);
            sdyj[I] = (j > iymin) ? t + sdyj[I] : t;
            i64 ix = gidx(&g, ic, jc, kc);
            jy[ix] = jy[ix] + sdyj[I];
          }
          if (k < izmax) {
            double t = wqz * dsz[K] * ((sx0[I] + 0.5 * dsx[I]) * sy0[J] +
                                        (0.5 * sx0[I] + ONETHIRD * dsx[I]) * dsy[J]);
            sdzk[I][J] = (k > izmin) ? t + sdzk[I][J] : t;
            i64 ix = gidx(&g, ic, jc, kc);
            jz[ix] = jz[ix] + sdzk[I][J];
          }
        }
      }
    }
  }
  return bad;
}

/* ------------------------------------------------------------------------
 * Generic-order routines: one shape order per axis.
 *
 * Weights of order n (0..3) at offset f from the base node, as the
 * generic routines write them (esirkepov_3d.F90:3588-3608,
 * charge_deposition_3d.F90:2160-2178, field_gathering_on_3d.F90:1170-1215):
 * order 0 is the single weight 1, order 3 the cubic with the constants
 * given (the gather's generic routines pass their own, below).          */
static inline void w_order(int n, double f, double *w, double sixth, double twothird) {
  switch (n) {
  case 0: w[0] = 1.0; break;
  case 1: w_lin(f, w); break;
  case 2: w_quad(f, w); break;
  default: {
    double o = 1.0 - f, f2 = f * f, o2 = o * o;
    w[0] = sixth * o2 * o;
    w[1] = twothird - f2 * (1.0 - f / 2.0);
    w[2] = twothird - o2 * (1.0 - o / 2.0);
    w[3] = sixth * f2 * f;
  }
  }
}

/* Node index of a shape of order n: nint for even orders, floor for odd
 * ones (esirkepov_3d.F90:3560-3574, charge_deposition_3d.F90:2128-2142). */
static inline i64 n_index(int n, double x) { return (n % 2 == 0) ? f_nint(x) : f_floor(x); }

/* ------------------------------------------------------------------------
 * Charge density, pxr_depose_rho_n (charge_deposition_3d.F90:2081-2228):
 * orders 0..3 on each axis, l_particles_weight (else every particle has
 * the weight w[0]) and l4symtry (|x|, |y|).  The first node of an axis is
 * j - n/2; rho(j+a, k+b, l+c) += sx(a) sy(b) sz(c) wq with wq = q w/vol.
 * rho is ADDED into.                                                    */
i64 orc_depose_rho_n(double *rho, i64 np, const double *xp, const double *yp,
                     const double *zp, const double *w, double q, double xmin,
                     double ymin, double zmin, double dx, double dy, double dz,
                     i64 nx, i64 ny, i64 nz, i64 ngx, i64 ngy, i64 ngz,
                     i64 nox, i64 noy, i64 noz, i64 l_particles_weight, i64 l4symtry) {
  if (nox < 0 || nox > 3 || noy < 0 || noy > 3 || noz < 0 || noz > 3) return -1;
  const int no[3] = {(int)nox, (int)noy, (int)noz};
  const double dxi = 1.0 / dx, dyi = 1.0 / dy, dzi = 1.0 / dz;
  const double invvol = dxi * dyi * dzi;
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  i64 bad = 0;
  for (i64 ip = 0; ip < np; ip++) {
    double p[3] = {(xp[ip] - xmin) * dxi, (yp[ip] - ymin) * dyi, (zp[ip] - zmin) * dzi};
    if (l4symtry) {
      p[0] = fabs(p[0]);
      p[1] = fabs(p[1]);
    }
    double s[3][4];
    i64 first[3];
    for (int a = 0; a < 3; a++) {
      i64 j = n_index(no[a], p[a]);
      w_order(no[a], p[a] - (double)j, s[a], ONESIXTH, TWOTHIRD);
      first[a] = j - no[a] / 2;
    }
    if (!ginside(&g, first[0], first[1], first[2]) ||
        !ginside(&g, first[0] + no[0], first[1] + no[1], first[2] + no[2])) {
      bad++;
      continue;
    }
    const double wq = l_particles_weight ? q * w[ip] * invvol : q * invvol * w[0];
    for (int c = 0; c <= no[2]; c++)
      for (int b = 0; b <= no[1]; b++)
        for (int a = 0; a <= no[0]; a++) {
          i64 ix = gidx(&g, first[0] + a, first[1] + b, first[2] + c);
          rho[ix] = rho[ix] + s[0][a] * s[1][b] * s[2][c] * wq;
        }
  }
  return bad;
}

/* pxr_depose_rho_n with nox = noy = noz = order (1..3), per-particle
 * weights, no symmetry: the node convention of the Esirkepov routines
 * above, so that rho and J share one shape (the continuity residual of
 * the tests). */
i64 orc_depose_rho(double *rho, i64 np, const double *xp, const double *yp,
                   const double *zp, const double *w, double q, double xmin,
                   double ymin, double zmin, double dx, double dy, double dz,
                   i64 nx, i64 ny, i64 nz, i64 ngx, i64 ngy, i64 ngz,
                   i64 order) {
  if (order < 1 || order > 3) return -1;
  return orc_depose_rho_n(rho, np, xp, yp, zp, w, q, xmin, ymin, zmin, dx, dy, dz, nx, ny, nz,
                          ngx, ngy, ngz, order, order, order, 1, 0);
}

/* ------------------------------------------------------------------------
 * pxr_depose_jxjyjz_esirkepov_n (esirkepov_3d.F90:3480-3836), as
 * depose_jxjyjz_generic calls it (l_particles_weight, no symmetry): orders
 * 0..3 on each axis.  Differences from the fixed-order routines that this
 * restatement keeps:
 *   - v = u * gaminv for every order, and the new position is formed again
 *     as xold + (dt/dx) v (the one-cell loop of :3553-3559), which may
 *     differ from the given position in the last bit;
 *   - the node index is nint for even orders (order 0 included).
 * The box of an axis is [min(0,d) - n/2, max(0,d) + (n+1)/2] around the
 * new node; particles that move more than one node are skipped (the
 * reference's work arrays do not hold them).                            */
i64 orc_depose_esirkepov_n(double *jx, double *jy, double *jz, i64 np,
                           const double *xp, const double *yp, const double *zp,
                           const double *uxp, const double *uyp, const double *uzp,
                           const double *gaminv, const double *w, double q, double xmin,
                           double ymin, double zmin, double dt, double dx, double dy,
                           double dz, i64 nx, i64 ny, i64 nz, i64 ngx, i64 ngy, i64 ngz,
                           i64 nox, i64 noy, i64 noz) {
  if (nox < 0 || nox > 3 || noy < 0 || noy > 3 || noz < 0 || noz > 3) return -1;
  const int no[3] = {(int)nox, (int)noy, (int)noz};
  const double dxi = 1.0 / dx, dyi = 1.0 / dy, dzi = 1.0 / dz;
  const double dts[3] = {dt * dxi, dt * dyi, dt * dzi};
  const double invdtdx = 1.0 / (dt * dy * dz);
  const double invdtdy = 1.0 / (dt * dx * dz);
  const double invdtdz = 1.0 / (dt * dx * dy);
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  i64 bad = 0;
  for (i64 ip = 0; ip < np; ip++) {
    const double p0[3] = {(xp[ip] - xmin) * dxi, (yp[ip] - ymin) * dyi, (zp[ip] - zmin) * dzi};
    const double v[3] = {uxp[ip] * gaminv[ip], uyp[ip] * gaminv[ip], uzp[ip] * gaminv[ip]};
    double s0[3][7], s1[3][7], ds[3][7];
    i64 base[3];
    int lo[3], hi[3], skip = 0;
    for (int a = 0; a < 3; a++) {
      const double xo = p0[a] - dts[a] * v[a];
      const double xn = xo + dts[a] * v[a];
      const int n = no[a], first = -(n / 2);
      double ww[4];
      memset(s0[a], 0, sizeof s0[a]);
      memset(s1[a], 0, sizeof s1[a]);
      const i64 i0 = n_index(n, xn);
      w_order(n, xn - (double)i0, ww, ONESIXTH, TWOTHIRD);
      for (int r = 0; r <= n; r++) s0[a][EOFF + first + r] = ww[r];
      const i64 i1 = n_index(n, xo);
      const int d = (int)(i1 - i0);
      if (d < -1 || d > 1) {
        skip = 1;
        break;
      }
      w_order(n, xo - (double)i1, ww, ONESIXTH, TWOTHIRD);
      for (int r = 0; r <= n; r++) s1[a][EOFF + first + d + r] = ww[r];
      for (int r = 0; r < 7; r++) ds[a][r] = s1[a][r] - s0[a][r];
      base[a] = i0;
      lo[a] = (d < 0 ? d : 0) - n / 2;
      hi[a] = (d > 0 ? d : 0) + (n + 1) / 2;
    }
    if (skip || !ginside(&g, base[0] + lo[0], base[1] + lo[1], base[2] + lo[2]) ||
        !ginside(&g, base[0] + hi[0], base[1] + hi[1], base[2] + hi[2])) {
      bad++;
      continue;
    }
    const double wq = q * w[ip];
    const double wqx = wq * invdtdx, wqy = wq * invdtdy, wqz = wq * invdtdz;
    const double *sx0 = s0[0], *sy0 = s0[1], *sz0 = s0[2];
    const double *dsx = ds[0], *dsy = ds[1], *dsz = ds[2];
    double sdyj[7] = {0}, sdzk[7][7] = {{0}};
    for (int k = lo[2]; k <= hi[2]; k++) {
      for (int j = lo[1]; j <= hi[1]; j++) {
        double sdxi = 0.0;
        for (int i = lo[0]; i <= hi[0]; i++) {
          i64 ic = base[0] + i, jc = base[1] + j, kc = base[2] + k;
          const int I = i + EOFF, J = j + EOFF, K = k + EOFF;
          if (i < hi[0]) {
            double t = wqx * dsx[I] * ((sy0[J] + 0.5 * dsy[J]) * sz0[K] +
                                        (0.5 * sy0[J] + ONETHIRD * dsy[J]) * dsz[K]);
            sdxi = (i > lo[0]) ? t + sdxi : t;
            i64 ix = gidx(&g, ic, jc, kc);
            jx[ix] = jx[ix] + sdxi;
          }
          if (j < hi[1]) {
            double t = wqy * dsy[J] * ((sz0[K] + 0.5 * dsz[K]) * sx0[I] +
                                        (0.5 * sz0[K] + ONETHIRD * dsz[K]) * dsx[I]);
            sdyj[I] = (j > lo[1]) ? t + sdyj[I] : t;
            i64 ix = gidx(&g, ic, jc, kc);
            jy[ix] = jy[ix] + sdyj[I];
          }
          if (k < hi[2]) {
            double t = wqz * dsz[K] * ((sx0[I] + 0.5 * dsx[I]) * sy0[J] +
                                        (0.5 * sx0[I] + ONETHIRD * dsx[I]) * dsy[J]);
            sdzk[I][J] = (k > lo[2]) ? t + sdzk[I][J] : t;
            i64 ix = gidx(&g, ic, jc, kc);
            jz[ix] = jz[ix] + sdzk[I][J];
          }
        }
      }
    }
  }
  return bad;
}

/* ------------------------------------------------------------------------
 * pxr_gete3d_n_energy_conserving and pxr_getb3d_n_energy_conserving
 * (field_gathering_on_3d.F90:1042-1365, 697-1026): orders 1..3 on each
 * axis, with l_lower_order_in_v, l_nodal and l4symtry.  Their constants
 * onesixth = 1./6. and twothird = 2./3. are default-real (single
 * precision) literals widened to FP64, so the cubic weights differ from
 * the order-3 routines' in the eighth digit; the restatement keeps them.
 * Per axis: the node-centred shape of order n at j (nint for even n,
 * floor for odd), first node j - n/2; the staggered shape of order
 * m = n - 1 (lower order in v) or n at j0, with
 *   lower: j0 = floor(x - s) for even n, floor(x + 1/2 - s) for odd n;
 *   else:  j0 = floor(x + 1/2 - s) for even n, floor(x - s) for odd n;
 * offset (x - s) - j0 and first node j0 - (n-1)/2 (lower) or j0 - n/2.
 * With l4symtry, x < 0 and y < 0 are mirrored and Ex, Ey, Bx, By take the
 * mirror's sign.  Results are ADDED into ex..bz (E first, then B).      */
static const double GN_SIXTH = (double)(1.0f / 6.0f);
static const double GN_TWOTHIRD = (double)(2.0f / 3.0f);

i64 orc_geteb3d_n(i64 np, const double *xp, const double *yp, const double *zp, double *ex,
                  double *ey, double *ez, double *bx, double *by, double *bz, double xmin,
                  double ymin, double zmin, double dx, double dy, double dz, i64 nx, i64 ny,
                  i64 nz, i64 ngx, i64 ngy, i64 ngz, i64 nox, i64 noy, i64 noz,
                  const double *exg, const double *eyg, const double *ezg, const double *bxg,
                  const double *byg, const double *bzg, i64 l4symtry, i64 l_lower_order_in_v,
                  i64 l_nodal) {
  if (nox < 1 || nox > 3 || noy < 1 || noy > 3 || noz < 1 || noz > 3) return -1;
  const int no[3] = {(int)nox, (int)noy, (int)noz};
  const int lower = l_lower_order_in_v != 0;
  const double s = l_nodal ? 0.0 : 0.5;
  const double dxi = 1.0 / dx, dyi = 1.0 / dy, dzi = 1.0 / dz;
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  i64 bad = 0;
  for (i64 ip = 0; ip < np; ip++) {
    double p[3] = {(xp[ip] - xmin) * dxi, (yp[ip] - ymin) * dyi, (zp[ip] - zmin) * dzi};
    double sgn[2] = {1.0, 1.0};
    if (l4symtry) {
      for (int a = 0; a < 2; a++)
        if (p[a] < 0.0) {
          p[a] = -p[a];
          sgn[a] = -1.0;
        }
    }
    double sn[3][4], ss[3][4];
    i64 fn[3], fs[3];
    int m[3];
    for (int a = 0; a < 3; a++) {
      const int n = no[a];
      const int even = (n % 2 == 0);
      const i64 j = even ? f_nint(p[a]) : f_floor(p[a]);
      const i64 j0 = (even == lower) ? f_floor(p[a] - s) : f_floor((p[a] + 0.5) - s);
      m[a] = lower ? n - 1 : n;
      w_order(n, p[a] - (double)j, sn[a], GN_SIXTH, GN_TWOTHIRD);
      w_order(m[a], (p[a] - s) - (double)j0, ss[a], GN_SIXTH, GN_TWOTHIRD);
      fn[a] = j - n / 2;
      fs[a] = j0 - (lower ? (n - 1) / 2 : n / 2);
    }
    i64 lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
      lo[a] = fn[a] < fs[a] ? fn[a] : fs[a];
      hi[a] = (fn[a] + no[a] > fs[a] + m[a]) ? fn[a] + no[a] : fs[a] + m[a];
    }
    if (!ginside(&g, lo[0], lo[1], lo[2]) || !ginside(&g, hi[0], hi[1], hi[2])) {
      bad++;
      continue;
    }
    const double *sx = sn[0], *sy = sn[1], *sz = sn[2];
    const double *sx0 = ss[0], *sy0 = ss[1], *sz0 = ss[2];
    const i64 j = fn[0], k = fn[1], l = fn[2], j0 = fs[0], k0 = fs[1], l0 = fs[2];
    for (int c = 0; c <= no[2]; c++)
      for (int b = 0; b <= no[1]; b++)
        for (int a = 0; a <= m[0]; a++)
          ex[ip] = ex[ip] + sx0[a] * sy[b] * sz[c] * exg[gidx(&g, j0 + a, k + b, l + c)] * sgn[0];
    for (int c = 0; c <= no[2]; c++)
      for (int b = 0; b <= m[1]; b++)
        for (int a = 0; a <= no[0]; a++)
          ey[ip] = ey[ip] + sx[a] * sy0[b] * sz[c] * eyg[gidx(&g, j + a, k0 + b, l + c)] * sgn[1];
    for (int c = 0; c <= m[2]; c++)
      for (int b = 0; b <= no[1]; b++)
        for (int a = 0; a <= no[0]; a++)
          ez[ip] = ez[ip] + sx[a] * sy[b] * sz0[c] * ezg[gidx(&g, j + a, k + b, l0 + c)];
    for (int c = 0; c <= m[2]; c++)
      for (int b = 0; b <= m[1]; b++)
        for (int a = 0; a <= no[0]; a++)
          bx[ip] = bx[ip] + sx[a] * sy0[b] * sz0[c] * bxg[gidx(&g, j + a, k0 + b, l0 + c)] * sgn[0];
    for (int c = 0; c <= m[2]; c++)
      for (int b = 0; b <= no[1]; b++)
        for (int a = 0; a <= m[0]; a++)
          by[ip] = by[ip] + sx0[a] * sy[b] * sz0[c] * byg[gidx(&g, j0 + a, k + b, l0 + c)] * sgn[1];
    for (int c = 0; c <= no[2]; c++)
      for (int b = 0; b <= m[1]; b++)
        for (int a = 0; a <= m[0]; a++)
          bz[ip] = bz[ip] + sx0[a] * sy0[b] * sz[c] * bzg[gidx(&g, j0 + a, k0 + b, l + c)];
  }
  return bad;
}

/* ------------------------------------------------------------------------
 * Yee FDTD, yee.F90:637-833 (pxrpush_em3d_evec) and :1258-1370
 * (pxrpush_em3d_bvec).  lo/hi are the inclusive loop bounds (the managers
 * pass -nxs..nx+nxs, maxwell_solver_manager.F90:60-70,412-426).          */
void orc_push_em3d_evec(i64 xlo, i64 xhi, i64 ylo, i64 yhi, i64 zlo, i64 zhi,
                        double *ex, double *ey, double *ez, const double *bx,
                        const double *by, const double *bz, const double *jx,
                        const double *jy, const double *jz, double mudt,
                        double dtsdx, double dtsdy, double dtsdz, i64 nx,
                        i64 ny, i64 nz, i64 ngx, i64 ngy, i64 ngz) {
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  const i64 sy = g.NX, sz = g.NX * g.NY;
  for (i64 l = zlo; l <= zhi; l++)
    for (i64 k = ylo; k <= yhi; k++)
      for (i64 j = xlo; j <= xhi; j++) {
        i64 c = gidx(&g, j, k, l);
        ex[c] = ex[c] + dtsdy * (bz[c] - bz[c - sy]) - dtsdz * (by[c] - by[c - sz]) - mudt * jx[c];
      }
  for (i64 l = zlo; l <= zhi; l++)
    for (i64 k = ylo; k <= yhi; k++)
      for (i64 j = xlo; j <= xhi; j++) {
        i64 c = gidx(&g, j, k, l);
        ey[c] = ey[c] - dtsdx * (bz[c] - bz[c - 1]) + dtsdz * (bx[c] - bx[c - sz]) - mudt * jy[c];
      }
  for (i64 l = zlo; l <= zhi; l++)
    for (i64 k = ylo; k <= yhi; k++)
      for (i64 j = xlo; j <= xhi; j++) {
        i64 c = gidx(&g, j, k, l);
        ez[c] = ez[c] + dtsdx * (by[c] - by[c - 1]) - dtsdy * (bx[c] - bx[c - sy]) - mudt * jz[c];
      }
}

void orc_push_em3d_bvec(i64 xlo, i64 xhi, i64 ylo, i64 yhi, i64 zlo, i64 zhi,
                        const double *ex, const double *ey, const double *ez,
                        double *bx, double *by, double *bz, double dtsdx,
                        double dtsdy, double dtsdz, i64 nx, i64 ny, i64 nz,
                        i64 ngx, i64 ngy, i64 ngz) {
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  const i64 sy = g.NX, sz = g.NX * g.NY;
  for (i64 l = zlo; l <= zhi; l++)
    for (i64 k = ylo; k <= yhi; k++)
      for (i64 j = xlo; j <= xhi; j++) {
        i64 c = gidx(&g, j, k, l);
        bx[c] = bx[c] - dtsdy * (ez[c + sy] - ez[c]) + dtsdz * (ey[c + sz] - ey[c]);
      }
  for (i64 l = zlo; l <= zhi; l++)
    for (i64 k = ylo; k <= yhi; k++)
      for (i64 j = xlo; j <= xhi; j++) {
        i64 c = gidx(&g, j, k, l);
        by[c] = by[c] + dtsdx * (ez[c + 1] - ez[c]) - dtsdz * (ex[c + sz] - ex[c]);
      }
  for (i64 l = zlo; l <= zhi; l++)
    for (i64 k = ylo; k <= yhi; k++)
      for (i64 j = xlo; j <= xhi; j++) {
        i64 c = gidx(&g, j, k, l);
        bz[c] = bz[c] - dtsdx * (ey[c + 1] - ey[c]) + dtsdy * (ex[c + sy] - ex[c]);
      }
}

/* Arbitrary-order FDTD, yee.F90:46-133 (pxrpush_em3d_evec_norder) and
 * :752-833 (pxrpush_em3d_bvec_norder).  dtsd* hold norder/2 coefficients
 * already scaled by c^2*dt/dx (E) or dt/(2dx) (B).  The stencil is cut
 * where it would leave the guard region (yee.F90:74,113).              */
static inline i64 mini(i64 a, i64 b) { return a < b ? a : b; }

void orc_push_em3d_evec_norder(double *ex, double *ey, double *ez,
                               const double *bx, const double *by,
                               const double *bz, const double *jx,
                               const double *jy, const double *jz, double mudt,
                               const double *dtsdx, const double *dtsdy,
                               const double *dtsdz, i64 nx, i64 ny, i64 nz,
                               i64 norderx, i64 nordery, i64 norderz, i64 ngx,
                               i64 ngy, i64 ngz, i64 nxs, i64 nys, i64 nzs,
                               i64 l_nodal) {
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  const i64 ist = l_nodal ? 0 : 1;
  for (i64 l = -nzs; l <= nz + nzs; l++)
    for (i64 k = -nys; k <= ny + nys; k++)
      for (i64 j = -nxs; j <= nx + nxs; j++) {
        i64 c = gidx(&g, j, k, l);
        double v = ex[c] - mudt * jx[c];
        i64 ny2 = mini(mini(nordery / 2, (ny - k) + ngy), k + ngy);
        for (i64 i = 1; i <= ny2; i++)
          v = v + dtsdy[i - 1] * (bz[gidx(&g, j, k + i - ist, l)] - bz[gidx(&g, j, k - i, l)]);
        i64 nz2 = mini(mini(norderz / 2, (nz - l) + ngz), l + ngz);
        for (i64 i = 1; i <= nz2; i++)
          v = v - dtsdz[i - 1] * (by[gidx(&g, j, k, l + i - ist)] - by[gidx(&g, j, k, l - i)]);
        ex[c] = v;
      }
  for (i64 l = -nzs; l <= nz + nzs; l++)
    for (i64 k = -nys; k <= ny + nys; k++)
      for (i64 j = -nxs; j <= nx + nxs; j++) {
        i64 c = gidx(&g, j, k, l);
        double v = ey[c] - mudt * jy[c];
        i64 nx2 = mini(mini(norderx / 2, (nx - j) + ngx), j + ngx);
        for (i64 i = 1; i <= nx2; i++)
          v = v - dtsdx[i - 1] * (bz[gidx(&g, j + i - ist, k, l)] - bz[gidx(&g, j - i, k, l)]);
        i64 nz2 = mini(mini(norderz / 2, (nz - l) + ngz), l + ngz);
        for (i64 i = 1; i <= nz2; i++)
          v = v + dtsdz[i - 1] * (bx[gidx(&g, j, k, l + i - ist)] - bx[gidx(&g, j, k, l - i)]);
        ey[c] = v;
      }
  for (i64 l = -nzs; l <= nz + nzs; l++)
    for (i64 k = -nys; k <= ny + nys; k++)
      for (i64 j = -nxs; j <= nx + nxs; j++) {
        i64 c = gidx(&g, j, k, l);
        double v = ez[c] - mudt * jz[c];
        i64 nx2 = mini(mini(norderx / 2, (nx - j) + ngx), j + ngx);
        for (i64 i = 1; i <= nx2; i++)
          v = v + dtsdx[i - 1] * (by[gidx(&g, j + i - ist, k, l)] - by[gidx(&g, j - i, k, l)]);
        i64 ny2 = mini(mini(nordery / 2, (ny - k) + ngy), k + ngy);
        for (i64 i = 1; i <= ny2; i++)
          v = v - dtsdy[i - 1] * (bx[gidx(&g, j, k + i - ist, l)] - bx[gidx(&g, j, k - i, l)]);
        ez[c] = v;
      }
}

void orc_push_em3d_bvec_norder(const double *ex, const double *ey,
                               const double *ez, double *bx, double *by,
                               double *bz, const double *dtsdx,
                               const double *dtsdy, const double *dtsdz,
                               i64 nx, i64 ny, i64 nz, i64 norderx,
                               i64 nordery, i64 norderz, i64 ngx, i64 ngy,
                               i64 ngz, i64 nxs, i64 nys, i64 nzs,
                               i64 l_nodal) {
  grid_t g = mkgrid(nx, ny, nz, ngx, ngy, ngz);
  const i64 ist = l_nodal ? 0 : 1;
  for (i64 l = -nzs; l <= nz + nzs; l++)
    for (i64 k = -nys; k <= ny + nys; k++)
      for (i64 j = -nxs; j <= nx + nxs; j++) {
        i64 c = gidx(&g, j, k, l);
        double v = bx[c];
        i64 ny2 = mini(mini(nordery / 2, (ny - k) + ngy), k + ngy);
        for (i64 i = 1; i <= ny2; i++)
          v = v - dtsdy[i - 1] * (ez[gidx(&g, j, k + i, l)] - ez[gidx(&g, j, k - i + ist, l)]);
        i64 nz2 = mini(mini(norderz / 2, (nz - l) + ngz), l + ngz);
        for (i64 i = 1; i <= nz2; i++)
          v = v + dtsdz[i - 1] * (ey[gidx(&g, j, k, l + i)] - ey[gidx(&g, j, k, l - i + ist)]);
        bx[c] = v;
      }
  for (i64 l = -nzs; l <= nz + nzs; l++)
    for (i64 k = -nys; k <= ny + nys; k++)
      for (i64 j = -nxs; j <= nx + nxs; j++) {
        i64 c = gidx(&g, j, k, l);
        double v = by[c];
        i64 nx2 = mini(mini(norderx / 2, (nx - j) + ngx), j + ngx);
        for (i64 i = 1; i <= nx2; i++)
          v = v + dtsdx[i - 1] * (ez[gidx(&g, j + i, k, l)] - ez[gidx(&g, j - i + ist, k, l)]);
        i64 nz2 = mini(mini(norderz / 2, (nz - l) + ngz), l + ngz);
        for (i64 i = 1; i <= nz2; i++)
          v = v - dtsdz[i - 1] * (ex[gidx(&g, j, k, l + i)] - ex[gidx(&g, j, k, l - i + ist)]);
        by[c] = v;
      }
  for (i64 l = -nzs; l <= nz + nzs; l++)
    for (i64 k = -nys; k <= ny + nys; k++)
      for (i64 j = -nxs; j <= nx + nxs; j++) {
        i64 c = gidx(&g, j, k, l);
        double v = bz[c];
        i64 nx2 = mini(mini(norderx / 2, (nx - j) + ngx), j + ngx);
        for (i64 i = 1; i <= nx2; i++)
          v = v - dtsdx[i - 1] * (ey[gidx(&g, j + i, k, l)] - ey[gidx(&g, j - i + ist, k, l)]);
        i64 ny2 = mini(mini(nordery / 2, (ny - k) + ngy), k + ngy);
        for (i64 i = 1; i <= ny2; i++)
          v = v + dtsdy[i - 1] * (ex[gidx(&g, j, k + i, l)] - ex[gidx(&g, j, k - i + ist, l)]);
        bz[c] = v;
      }
}

/* Finite-difference weights for a staggered (or nodal) first derivative,
 * Fornberg's recursion as used by FD_weights (submain.F90:927-984).
 * coeffs receives norder/2 values.                                       */
void orc_fd_weights(double *coeffs, i64 norder, i64 l_nodal) {
  const double z = l_nodal ? 0.0 : 0.5;
  const i64 mdeg = 1, n = norder + 1;
  double *xg = (double *)malloc(sizeof(double) * n);
  double *c = (double *)calloc((size_t)(mdeg + 1) * n, sizeof(double));
#define C(k, i) c[(k) + (mdeg + 1) * (i)]
  for (i64 i = 0; i < n; i++) xg[i] = (double)(i - n / 2 + 1);
  double c1 = 1.0, c4 = xg[0] - z;
  C(0, 0) = 1.0;
  for (i64 i = 1; i < n; i++) {
    i64 mn = (i + 1 < mdeg + 1) ? i + 1 : mdeg + 1;
    double c2 = 1.0, c5 = c4;
    c4 = xg[i] - z;
    for (i64 j = 0; j < i; j++) {
      double c3 = xg[i] - xg[j];
      c2 = c2 * c3;
      /* The reference updates column j only when j == i-1 (the column
       * update sits inside the IF at submain.F90:964-973). */
      if (j == i - 1) {
        for (i64 k = 1; k < mn; k++) C(k, i) = c1 * ((double)k * C(k - 1, i - 1) - c5 * C(k, i - 1)) / c2;
        C(0, i) = -c1 * c5 * C(0, i - 1) / c2;
        for (i64 k = 1; k < mn; k++) C(k, j) = (c4 * C(k, j) - (double)k * C(k - 1, j)) / c3;
        C(0, j) = c4 * C(0, j) / c3;
      }
    }
    c1 = c2;
  }
  for (i64 i = 1; i <= norder / 2; i++) coeffs[i - 1] = C(mdeg, norder / 2 + i - 1);
#undef C
  free(xg);
  free(c);
}

/* ------------------------------------------------------------------------
 * Stable counting sort of one tile's particles by sorting-grid cell,
 * pxr_particle_bin_sorting (sorting.F90:247-361).  Cell key (1-based):
 * iz*nx3*ny3 + iy*nx3 + ix + 1 with nx3 = ceiling((xmax-xmin)/dxf).  Only
 * the first pid column (np2 values) is permuted (sorting.F90:264).
 * Returns the number of particles whose key falls outside [1, nbhc]
 * (the arrays are then left unchanged; the reference indexes out of bounds).
 * If keys_out is non-NULL, it receives each particle's key (input order). */
i64 orc_particle_bin_sorting(i64 np2, double *xp, double *yp, double *zp,
                             double *ux, double *uy, double *uz, double *gam,
                             double *pid, double xmin2, double ymin2,
                             double zmin2, double xmax2, double ymax2,
                             double zmax2, double dxf, double dyf, double dzf,
                             i64 *keys_out) {
  const double dxi = 1.0 / dxf, dyi = 1.0 / dyf, dzi = 1.0 / dzf;
  const i64 nx3 = (i64)ceil((xmax2 - xmin2) * dxi);
  const i64 ny3 = (i64)ceil((ymax2 - ymin2) * dyi);
  const i64 nz3 = (i64)ceil((zmax2 - zmin2) * dzi);
  const i64 nbhc = nx3 * ny3 * nz3;
  if (np2 <= 0) return 0;
  i64 *key = (i64 *)malloc(sizeof(i64) * np2);
  i64 bad = 0;
  for (i64 ip = 0; ip < np2; ip++) {
    i64 ix = f_floor((xp[ip] - xmin2) * dxi);
    i64 iy = f_floor((yp[ip] - ymin2) * dyi);
    i64 iz = f_floor((zp[ip] - zmin2) * dzi);
    key[ip] = iz * nx3 * ny3 + iy * nx3 + ix + 1;
    if (key[ip] < 1 || key[ip] > nbhc) bad++;
    if (keys_out) keys_out[ip] = key[ip];
  }
  if (bad) {
    free(key);
    return bad;
  }
  i64 *start = (i64 *)calloc((size_t)nbhc + 1, sizeof(i64));
  for (i64 ip = 0; ip < np2; ip++) start[key[ip]]++;
  i64 acc = 0;
  for (i64 c = 1; c <= nbhc; c++) {
    i64 t = start[c];
    start[c] = acc;
    acc += t;
  }
  double *tmp = (double *)malloc(sizeof(double) * np2);
  double *arrs[8] = {xp, yp, zp, ux, uy, uz, gam, pid};
  i64 *dest = (i64 *)malloc(sizeof(i64) * np2);
  for (i64 ip = 0; ip < np2; ip++) dest[ip] = start[key[ip]]++;
  for (int a = 0; a < 8; a++) {
    for (i64 ip = 0; ip < np2; ip++) tmp[dest[ip]] = arrs[a][ip];
    memcpy(arrs[a], tmp, sizeof(double) * np2);
  }
  free(dest);
  free(tmp);
  free(start);
  free(key);
  return 0;
}

/* ------------------------------------------------------------------------
 * Periodic guard handling on one rank, where every MPI neighbour is the
 * rank itself.
 * field_bc -> exchange_mpi_3d_grid_array_with_guards
 * (field_boundaries.F90:107-260): per axis, nodes n..n+g := nodes 0..g,
 * then nodes -g..0 := nodes n-g..n; x, then y (all x), then z.           */
void orc_field_bc_periodic(double *f, i64 nx, i64 ny, i64 nz, i64 gx, i64 gy,
                           i64 gz) {
  grid_t g = mkgrid(nx, ny, nz, gx, gy, gz);
  for (i64 k = -gz; k <= nz + gz; k++)
    for (i64 j = -gy; j <= ny + gy; j++) {
      for (i64 i = 0; i <= gx; i++) f[gidx(&g, nx + i, j, k)] = f[gidx(&g, i, j, k)];
      for (i64 i = 0; i <= gx; i++) f[gidx(&g, -gx + i, j, k)] = f[gidx(&g, nx - gx + i, j, k)];
    }
  for (i64 k = -gz; k <= nz + gz; k++)
    for (i64 i = -gx; i <= nx + gx; i++) {
      for (i64 j = 0; j <= gy; j++) f[gidx(&g, i, ny + j, k)] = f[gidx(&g, i, j, k)];
      for (i64 j = 0; j <= gy; j++) f[gidx(&g, i, -gy + j, k)] = f[gidx(&g, i, ny - gy + j, k)];
    }
  for (i64 j = -gy; j <= ny + gy; j++)
    for (i64 i = -gx; i <= nx + gx; i++) {
      for (i64 k = 0; k <= gz; k++) f[gidx(&g, i, j, nz + k)] = f[gidx(&g, i, j, k)];
      for (i64 k = 0; k <= gz; k++) f[gidx(&g, i, j, -gz + k)] = f[gidx(&g, i, j, nz - gz + k)];
    }
}

/* summation_bcs (field_boundaries.F90:403-500) on one periodic rank: per
 * axis, a(0:g) += a(n:n+g); then a(n-g:n) += a(-g:0) using the values just
 * updated (so node n receives a(n) + a(0) + a(n); the E update at node n
 * is later overwritten by field_bc).                                     */
void orc_current_bc_periodic(double *a, i64 nx, i64 ny, i64 nz, i64 gx, i64 gy,
                             i64 gz) {
  grid_t g = mkgrid(nx, ny, nz, gx, gy, gz);
  for (i64 k = -gz; k <= nz + gz; k++)
    for (i64 j = -gy; j <= ny + gy; j++) {
      for (i64 i = 0; i <= gx; i++) a[gidx(&g, i, j, k)] += a[gidx(&g, nx + i, j, k)];
      for (i64 i = 0; i <= gx; i++) a[gidx(&g, nx - gx + i, j, k)] += a[gidx(&g, -gx + i, j, k)];
    }
  for (i64 k = -gz; k <= nz + gz; k++)
    for (i64 i = -gx; i <= nx + gx; i++) {
      for (i64 j = 0; j <= gy; j++) a[gidx(&g, i, j, k)] += a[gidx(&g, i, ny + j, k)];
      for (i64 j = 0; j <= gy; j++) a[gidx(&g, i, ny - gy + j, k)] += a[gidx(&g, i, -gy + j, k)];
    }
  for (i64 j = -gy; j <= ny + gy; j++)
    for (i64 i = -gx; i <= nx + gx; i++) {
      for (i64 k = 0; k <= gz; k++) a[gidx(&g, i, j, k)] += a[gidx(&g, i, j, nz + k)];
      for (i64 k = 0; k <= gz; k++) a[gidx(&g, i, j, nz - gz + k)] += a[gidx(&g, i, j, -gz + k)];
    }
}

/* ---- divergence diagnostics (diags.F90) -------------------------------- */
/* calc_field_div (diags.F90:129-156): backward differences of E at the
 * nodes 0..n of each axis, summed as (x + y) + z.                         */
void orc_calc_field_div(double *dv, const double *ex, const double *ey, const double *ez,
                        i64 nx, i64 ny, i64 nz, i64 gx, i64 gy, i64 gz, double dx,
                        double dy, double dz) {
  grid_t g = mkgrid(nx, ny, nz, gx, gy, gz);
  const double ix = 1.0 / dx, iy = 1.0 / dy, iz = 1.0 / dz;
  for (i64 l = 0; l <= nz; l++)
    for (i64 k = 0; k <= ny; k++)
      for (i64 j = 0; j <= nx; j++)
        dv[gidx(&g, j, k, l)] = ix * (ex[gidx(&g, j, k, l)] - ex[gidx(&g, j - 1, k, l)]) +
                                iy * (ey[gidx(&g, j, k, l)] - ey[gidx(&g, j, k - 1, l)]) +
                                iz * (ez[gidx(&g, j, k, l)] - ez[gidx(&g, j, k, l - 1)]);
}

/* calc_field_divB (diags.F90:170-197): forward differences of B at the
 * nodes 1..n of each axis.                                                 */
void orc_calc_field_divb(double *dv, const double *bx, const double *by, const double *bz,
                         i64 nx, i64 ny, i64 nz, i64 gx, i64 gy, i64 gz, double dx,
                         double dy, double dz) {
  grid_t g = mkgrid(nx, ny, nz, gx, gy, gz);
  const double ix = 1.0 / dx, iy = 1.0 / dy, iz = 1.0 / dz;
  for (i64 l = 1; l <= nz; l++)
    for (i64 k = 1; k <= ny; k++)
      for (i64 j = 1; j <= nx; j++)
        dv[gidx(&g, j, k, l)] = ix * (bx[gidx(&g, j + 1, k, l)] - bx[gidx(&g, j, k, l)]) +
                                iy * (by[gidx(&g, j, k + 1, l)] - by[gidx(&g, j, k, l)]) +
                                iz * (bz[gidx(&g, j, k, l + 1)] - bz[gidx(&g, j, k, l)]);
}

/* get_loc_norm_divErho / get_loc_norm_2 (diags.F90:1222-1293): the sum over
 * the nodes 1..n of each axis of (a*eps0 - b)^2 (eps0 = 1, normalised
 * units), or of a^2 when b is NULL; summed in the reference's loop order. */
double orc_node_norm2(const double *a, const double *b, i64 nx, i64 ny, i64 nz, i64 gx, i64 gy,
                      i64 gz) {
  grid_t g = mkgrid(nx, ny, nz, gx, gy, gz);
  double s = 0.0;
  for (i64 l = 1; l <= nz; l++)
    for (i64 k = 1; k <= ny; k++)
      for (i64 j = 1; j <= nx; j++) {
        const i64 p = gidx(&g, j, k, l);
        const double v = b ? a[p] * 1.0 - b[p] : a[p];
        s = s + v * v;
      }
  return s;
}

/* ---- laser antenna (laser_pusher_manager_3d.F90) ---------------------- */
/* Complex numbers as (re, im) pairs, in the reference's evaluation order;
 * the division is Smith's (1962) algorithm.                              */
typedef struct { double re, im; } cplx;
static inline cplx c_mul(cplx a, cplx b) {
  cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static inline cplx c_scale(cplx a, double s) { cplx r = {a.re * s, a.im * s}; return r; }
static inline cplx c_sub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx c_add(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx c_div(cplx a, cplx b) {
  cplx r;
  if (fabs(b.re) >= fabs(b.im)) {
    const double t = b.im / b.re, den = b.re + b.im * t;
    r.re = (a.re + a.im * t) / den;
    r.im = (a.im - a.re * t) / den;
  } else {
    const double t = b.re / b.im, den = b.im + b.re * t;
    r.re = (a.re * t + a.im) / den;
    r.im = (a.im * t - a.re) / den;
  }
  return r;
}
static inline cplx c_exp(cplx a) {
  const double e = exp(a.re);
  cplx r = {e * cos(a.im), e * sin(a.im)};
  return r;
}
/* x**n for an integer n >= 0 by binary powering (as compilers expand a
 * real**integer power) */
static inline double ipow(double x, i64 n) {
  double r = 1.0;
  while (n > 0) {
    if (n & 1) r *= x;
    x *= x;
    n >>= 1;
  }
  return r;
}

static const double PXR_PI = 3.14159265358979323;

/* gaussian_profile, laser_pusher_manager_3d.F90:368-415 (clight = 1) */
static void gaussian_amp(double xx, double yy, double *amp, double emax1, double emax2,
                         const double *pv1, const double *pv2, double k0, cplx qz,
                         double tau, double t, double tpeak, i64 torder, double polangle) {
  const cplx j = {0.0, 1.0};
  const cplx jk = c_scale(j, k0);
  const cplx a = c_scale(c_scale(jk, 1.0), t - tpeak);
  const cplx b = c_div(c_scale(jk, xx * xx + yy * yy), c_scale(qz, 2.0));
  cplx u1, u2;
  if (torder == 0) {
    u1 = c_sub(a, b);
    u2 = c_sub(a, b);
  } else {
    const double env = ipow((t - tpeak) / tau, torder);
    const cplx e = {env, 0.0};
    u1 = c_sub(c_sub(a, b), e);
    u2 = c_add(c_sub(c_sub(a, b), e), c_scale(j, polangle * 2.0 * PXR_PI));
  }
  u1 = c_scale(c_exp(u1), emax1);
  u2 = c_scale(c_exp(u2), emax2);
  for (int d = 0; d < 3; d++) amp[d] = c_add(c_scale(u1, pv1[d]), c_scale(u2, pv2[d])).im;
}

/* hanning_profile, laser_pusher_manager_3d.F90:426-478 (clight = 1) */
static void hanning_amp(double xx, double yy, double *amp, double emax1, double emax2,
                        const double *pv1, const double *pv2, double k0, cplx qz,
                        double t, double tpeak, double polangle) {
  if (!(t <= 2.0 * tpeak)) {
    amp[0] = amp[1] = amp[2] = 0.0;
    return;
  }
  const double idd = 1.0;
  const cplx j = {0.0, 1.0};
  const cplx jk = c_scale(j, k0);
  const cplx x = c_div((cplx){xx * xx + yy * yy, 0.0}, c_scale(qz, 2.0));
  const cplx arg = c_sub(x, (cplx){1.0 * (t - tpeak), 0.0});
  cplx u1 = c_scale(c_mul(jk, arg), -1.0);
  cplx u2 = c_add(c_scale(c_mul(jk, arg), -1.0),
                  c_scale(c_scale(c_scale(j, polangle), 2.0), PXR_PI));
  const double win = 0.5 - 0.5 * cos(t * 2.0 * PXR_PI / (2.0 * tpeak));
  u1 = c_scale(c_scale(c_scale(c_exp(u1), idd), emax1), win);
  u2 = c_scale(c_scale(c_scale(c_exp(u2), idd), emax2), win);
  for (int d = 0; d < 3; d++) amp[d] = c_add(c_scale(u1, pv1[d]), c_scale(u2, pv2[d])).re;
}

/* The antenna's field amplitude at antenna-frame coordinates (xx, yy):
 * window 0 = gaussian_profile, 1 = hanning_profile.  amp is (n, 3). */
void orc_laser_profile(i64 window, i64 n, const double *xx, const double *yy, double *amp,
                       double emax1, double emax2, const double *pv1, const double *pv2,
                       double k0, double qz_re, double qz_im, double tau, double t,
                       double tpeak, i64 torder, double polangle) {
  const cplx qz = {qz_re, qz_im};
  for (i64 i = 0; i < n; i++) {
    if (window == 0)
      gaussian_amp(xx[i], yy[i], amp + 3 * i, emax1, emax2, pv1, pv2, k0, qz, tau, t, tpeak,
                   torder, polangle);
    else
      hanning_amp(xx[i], yy[i], amp + 3 * i, emax1, emax2, pv1, pv2, k0, qz, t, tpeak,
                  polangle);
  }
}

/* laserp_pusher_gaussian / laserp_pusher_hanning,
 * laser_pusher_manager_3d.F90:195-271 / :281-356: the antenna-frame
 * coordinates are pid columns 2 and 3 (pid is np x npid, column-major);
 * u = charge * amp * (0.01 c / emax), gaminv = 1, x += dt * u.  (The
 * reference pushes with the module's dt, which its one caller passes as
 * dtt; dt here is that value.)                                           */
void orc_laserp_pusher(i64 window, i64 np, i64 npid, const double *pid, double *xp,
                       double *yp, double *zp, double *uxp, double *uyp, double *uzp,
                       double *gaminv, double dt, double emax, double emax1, double emax2,
                       const double *pv1, const double *pv2, double k0, double qz_re,
                       double qz_im, double charge, double tau, double t, double tpeak,
                       i64 torder, double polangle) {
  const cplx qz = {qz_re, qz_im};
  const double disp_max = 0.01 * 1.0;
  const double coeff = disp_max / emax;
  (void)npid;
  for (i64 i = 0; i < np; i++) {
    double amp[3];
    const double xx = pid[np + i], yy = pid[2 * np + i];
    if (window == 0)
      gaussian_amp(xx, yy, amp, emax1, emax2, pv1, pv2, k0, qz, tau, t, tpeak, torder,
                   polangle);
    else
      hanning_amp(xx, yy, amp, emax1, emax2, pv1, pv2, k0, qz, t, tpeak, polangle);
    uxp[i] = charge * amp[0] * coeff;
    uyp[i] = charge * amp[1] * coeff;
    uzp[i] = charge * amp[2] * coeff;
    gaminv[i] = 1.0;
    xp[i] = xp[i] + dt * uxp[i];
    yp[i] = yp[i] + dt * uyp[i];
    zp[i] = zp[i] + dt * uzp[i];
  }
}
