// pxr_leafbin.h -- arguments of the leaf deposition kernels (global
// atomics, pxr_leaf.hip) and of their binned form (pxr_leafbin.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "pxr_device.h"
#include "pxr_kernels.h"

namespace pxr {

// current deposition of one call of depose_jxjyjz
struct DepoArgs {
  i64 np;
  const double *xp, *yp, *zp, *ux, *uy, *uz, *gi, *w;
  double q, xmin, ymin, zmin, dt, dx, dy, dz, s;
  EsirkConst<double> ec;
  GridGeom g;
  double *jx, *jy, *jz;
  unsigned long long *bad;
};

// charge density of one call of pxr_depose_rho_n
struct RhoArgs {
  i64 np;
  const double *xp, *yp, *zp, *w;
  double q, xmin, ymin, zmin, dxi, dyi, dzi, invvol;
  int per_particle_weight, l4symtry;
  GridGeom g;
  double *rho;
  unsigned long long *bad;
};

// global-memory accumulator of the current (FP64 atomics), node (0,0,0)
// at p[c]
struct GlobalAcc {
  double *p[3];
  i64 sy, sz;
  PXR_DEV void add(int c, int i, int j, int k, double v) const {
    unsafeAtomicAdd(p[c] + (i + sy * j + sz * k), v);
  }
  PXR_DEV double *row(int c, int i, int j, int k) const { return p[c] + (i + sy * j + sz * k); }
  PXR_DEV void add_row(double *r, int o, double v) const { unsafeAtomicAdd(r + o, v); }
};

// Binned forms: the particles are counted into cubic blocks of nodes,
// ordered by block, and each block's current (or charge) is summed in LDS
// before one global add per node.  Return PXRHIP_OK, or an error code
// (scratch memory could not be allocated) before anything is launched.
int leaf_binned_esirkepov(const DepoArgs &a, int order, hipStream_t st);
int leaf_binned_rho(const RhoArgs &a, int order, hipStream_t st);
// whether a call of np particles on grid g takes the binned form: at least
// the runtime's leaf_bin_min particles and one per block, and no graph
// capture open (the scratch may have to grow)
bool leaf_binned_esirkepov_pays(i64 np, const GridGeom &g);
bool leaf_binned_rho_pays(i64 np, const GridGeom &g);

}  // namespace pxr
