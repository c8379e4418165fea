// pxr_generic.h -- arguments and launchers of the generic-order leaf kernels
// (pxr_generic.hip): one shape order per axis.
#pragma once
#include <hip/hip_runtime.h>

#include "pxr_device.h"

namespace pxr {

struct GatherNArgs {
  i64 np;
  const double *xp, *yp, *zp;
  double *ex, *ey, *ez, *bx, *by, *bz;
  double xmin, ymin, zmin, dxi, dyi, dzi, s;
  GridGeom g;
  const double *exg, *eyg, *ezg, *bxg, *byg, *bzg;
  int no[3];  // orders 1..3
  int lower, l4symtry;
  unsigned long long *bad;
};

struct EsirkNArgs {
  i64 np;
  const double *xp, *yp, *zp, *ux, *uy, *uz, *gi, *w;
  double q, xmin, ymin, zmin, dxi, dyi, dzi;
  double dts[3];    // dt/dx, dt/dy, dt/dz
  double invdt[3];  // 1/(dt dy dz), 1/(dt dx dz), 1/(dt dx dy)
  GridGeom g;
  double *jx, *jy, *jz;
  int no[3];  // orders 0..3
  unsigned long long *bad;
};

struct RhoNArgs {
  i64 np;
  const double *xp, *yp, *zp, *w;
  double q, xmin, ymin, zmin, dxi, dyi, dzi, invvol;
  int per_particle_weight, l4symtry;
  GridGeom g;
  double *rho;
  int no[3];  // orders 0..3
  unsigned long long *bad;
};

void launch_gather_n(const GatherNArgs &a, hipStream_t st);
void launch_esirkepov_n(const EsirkNArgs &a, hipStream_t st);
void launch_rho_n(const RhoNArgs &a, hipStream_t st);

}  // namespace pxr
