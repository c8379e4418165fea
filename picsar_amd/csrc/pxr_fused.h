// pxr_fused.h -- the fused gather + push + deposit kernel of the tiled step
// (shared by the FP64 and the mixed-precision instantiations, which are
// compiled in separate translation units).
//
// Work of field_gathering_plus_particle_pusher_cacheblock_sub
// (particle_pusher_manager_3d.F90:431-672) fused with the current deposition
// of the pushed particles (pxrdepose_currents_on_grid_jxjyjz,
// current_deposition_manager_3d.F90:224-1830), for one species over all its
// tiles.
#pragma once
#include <limits.h>

#include "pxr_device.h"
#include "pxr_kernels.h"
#include "pxr_mixed.h"

namespace pxr {

// Node blocks of a work item.  The particles of a work item sit in known
// cells of the sorting grid (tile t, plane z, rows y0..y1), so the index
// ranges of their gather stencils and of their currents follow from the
// positions bounding those cells; the index functions of the reference's
// shapes are non-decreasing in the position.
struct BoxGeom {
  double x0[3], di[3];   // grid origin and inverse cell size
  double sd[3];          // sorting cell size
  double mv[3];          // motion margin of the current block, in cells
  double s;              // staggering shift (1/2)
  int N, M, depo;        // shape orders (nodal, staggered) and deposition (0 Esirkepov, 3 direct)
  int glo[3], ghi[3];    // node range of the grid arrays
  const double *sorg;    // sorting-grid origin of every tile
  const double *tbox;    // membership [tmin, tmax) of every tile
  const int *n3;         // sorting cells of every tile
};

PXR_HD void gather_range(const BoxGeom &b, double ulo, double uhi, int &lo, int &hi) {
  const int N = b.N, M = b.M;
  auto jn = [&](double v) { return (N == 2 ? f_nint(v) : f_floor(v)) - (N == 1 ? 0 : 1); };
  auto js = [&](double v) {
    return (M == 0 || M == 2) ? f_floor((v + 0.5) - b.s) - (M == 2 ? 1 : 0)
                              : f_floor(v - b.s) - (M == 3 ? 1 : 0);
  };
  lo = min(jn(ulo), js(ulo));
  hi = max(jn(uhi) + N, js(uhi) + M);
}

PXR_HD void deposit_range(const BoxGeom &b, double ulo, double uhi, double m, int &lo, int &hi) {
  const int N = b.N;
  if (b.depo == 0) {  // Esirkepov: old and new node index, support ESup<N>
    auto ei = [&](double v) { return N == 2 ? f_nint(v) : f_floor(v); };
    lo = ei(ulo - m) + (N == 1 ? 0 : -1);
    hi = ei(uhi + m) + (N == 3 ? 2 : 1);
  } else {  // direct: nodal and staggered shapes at the mid position
    auto dj = [&](double v) { return N == 2 ? f_nint(v) - 1 : f_floor(v) - (N == 1 ? 0 : 1); };
    lo = min(dj(ulo - m), dj((ulo - m) - b.s));
    hi = max(dj(uhi + m), dj((uhi + m) - b.s)) + N;
  }
}

// normalized coordinate range of the particles of tile t in sorting rows
// y0..y1 of planes z0..z1 (intersected with the tile's membership)
PXR_DEV void rows_range(const BoxGeom &b, int t, int z0, int z1, int y0, int y1, double *ulo,
                        double *uhi) {
  const double *so = b.sorg + 3 * t, *tb = b.tbox + 6 * t;
  const int c0[3] = {0, y0, z0};
  const int c1[3] = {b.n3[3 * t] - 1, y1, z1};
  for (int d = 0; d < 3; d++) {
    const double plo = fmax(so[d] + b.sd[d] * c0[d], tb[d]);
    const double phi = fmin(so[d] + b.sd[d] * (c1[d] + 1), tb[3 + d]);
    ulo[d] = (plo - b.x0[d]) * b.di[d] - 1e-7;
    uhi[d] = (phi - b.x0[d]) * b.di[d] + 1e-7;
  }
}

// field block (FR0, fe) and current block (JR0, je) for that range, clamped
// to the arrays; returns the doubles they take in LDS (6 + 3 arrays)
PXR_HD i64 blocks_for(const BoxGeom &b, const double *ulo, const double *uhi, int *FR0, int *fe,
                       int *JR0, int *je) {
  for (int d = 0; d < 3; d++) {
    int lo, hi;
    gather_range(b, ulo[d], uhi[d], lo, hi);
    FR0[d] = max(lo, b.glo[d]);
    fe[d] = max(min(hi, b.ghi[d]) - FR0[d] + 1, 0);
    deposit_range(b, ulo[d], uhi[d], b.mv[d], lo, hi);
    JR0[d] = max(lo, b.glo[d]);
    je[d] = max(min(hi, b.ghi[d]) - JR0[d] + 1, 0);
  }
  return 6 * (i64)fe[0] * fe[1] * fe[2] + 3 * (i64)je[0] * je[1] * je[2];
}

struct Chunk {
  long long begin;
  int n;
  int tile;
  int z, y0, y1;  // sorting-grid planes z..z1 and rows y0..y1 of the particles (several
  int z1;         // planes: whole planes); z < 0: unknown
  int s;          // largest lane stride of the deposition (0: none), from its density
  int pad;
};

struct FusedArgs {
  const Chunk *chunks;
  const int *nchunks;
  int *queue;
  double *x, *y, *z;
  double *ux, *uy, *uz, *gi;  // FP64 species
  const double *w;
  float *fux, *fuy, *fuz, *fgi;  // mixed-precision species
  const float *fw;
  const double *F[6];
  double *J[3];
  GridGeom g;
  double xmin, ymin, zmin, dxi, dyi, dzi, s;
  double cst;            // Boris: q*dt/(2m); Vay: q*dt/m
  double dt;
  EsirkConst<double> ec;
  double q, dx, dy, dz;  // direct deposition
  unsigned long long *bad;
  BoxGeom bg;            // work-item block geometry
  int lds_bytes;         // dynamic LDS of a workgroup
  int j_bytes;           // largest current block kept in LDS
  int spread;            // lanes take particles S apart
  int sstride;           // largest S (0: spread over the whole work item)
  int jpad;              // odd pitches for the LDS current block
  int jacc;              // mixed precision: LDS current accumulation (see LdsAccMx)
  int *ovf_count;        // work items whose blocks do not fit in LDS ...
  int2 *ovf;             // ... (work item, stage): 0 untouched, 1 pushed
  int ovf_cap;
  int phase;             // PHASE_PUSH | PHASE_DEPOSIT
  double fxb;            // bound of |current contribution| per unit |q w| (fixed point)
  const int *order;      // work items in the order the queue hands them out (null: as built)
};

// phases of k_fused: gather + push + position push, current deposition.
// PHASE_CUT (fixed-point current of a mixed-precision species): a work item
// whose current box does not fit is pushed by k_fused and deposited by
// k_fused_cut through a block cut down to fit.
constexpr int PHASE_PUSH = 1, PHASE_DEPOSIT = 2, PHASE_CUT = 4;

struct Box6 {
  int lo[3], hi[3];
};

PXR_DEV void box_reset(Box6 &b) {
  for (int d = 0; d < 3; d++) {
    b.lo[d] = INT_MAX;
    b.hi[d] = INT_MIN;
  }
}

PXR_DEV void wave_minmax(Box6 &b) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      b.lo[d] = min(b.lo[d], __shfl_xor(b.lo[d], o));
      b.hi[d] = max(b.hi[d], __shfl_xor(b.hi[d], o));
    }
  }
}

// Current accumulators: add(c, i, j, k, v) adds v to node (i, j, k) of
// component c; row(c, i, j, k) is the address of that node, and
// add_row(row, r, v) adds v to the node r further along x.
struct LdsAcc {
  double *p[3];
  int i0, j0, k0, sy, sz;
  PXR_DEV void add(int c, int i, int j, int k, double v) const {
    atomicAdd(p[c] + ((i - i0) + sy * (j - j0) + sz * (k - k0)), v);
  }
  PXR_DEV double *row(int c, int i, int j, int k) const {
    return p[c] + ((i - i0) + sy * (j - j0) + sz * (k - k0));
  }
  PXR_DEV void add_row(double *r, int o, double v) const { atomicAdd(r + o, v); }
};

// The same in fixed point: v is added as fixed_point(v, scale) to 64-bit
// integer counters, so a block's sums are exact and do not depend on the
// order of the atomics.
struct LdsAccFx {
  unsigned long long *p[3];
  int i0, j0, k0, sy, sz;
  double scale;
  PXR_DEV void add(int c, int i, int j, int k, double v) const {
    atomicAdd(p[c] + ((i - i0) + sy * (j - j0) + sz * (k - k0)), fixed_point(v, scale));
  }
  PXR_DEV unsigned long long *row(int c, int i, int j, int k) const {
    return p[c] + ((i - i0) + sy * (j - j0) + sz * (k - k0));
  }
  PXR_DEV void add_row(unsigned long long *r, int o, double v) const {
    atomicAdd(r + o, fixed_point(v, scale));
  }
};

struct GAcc {
  double *p[3];
  GridGeom g;
  PXR_DEV void add(int c, int i, int j, int k, double v) const {
    unsafeAtomicAdd(p[c] + g.idx(i, j, k), v);
  }
  PXR_DEV double *row(int c, int i, int j, int k) const { return p[c] + g.idx(i, j, k); }
  PXR_DEV void add_row(double *r, int o, double v) const { unsafeAtomicAdd(r + o, v); }
};

PXR_DEV bool in_block(const int *lo, const int *hi, const int *R0, const int *e) {
  return lo[0] >= R0[0] && lo[1] >= R0[1] && lo[2] >= R0[2] && hi[0] < R0[0] + e[0] &&
         hi[1] < R0[1] + e[1] && hi[2] < R0[2] + e[2];
}

// Bound of a particle's largest current contribution, in units of a.fxb:
// |q w| times the largest of its displacement in cells (Esirkepov: the
// running sums of shape differences are bounded by it) or of its velocity
// components (direct deposition).
template <int DEPO>
PXR_DEV double fx_weight(const FusedArgs &a, double qw, double vx, double vy, double vz) {
  const double f = DEPO == 0 ? fmax(fabs(vx) * a.ec.dtsdx0,
                                    fmax(fabs(vy) * a.ec.dtsdy0, fabs(vz) * a.ec.dtsdz0))
                             : fmax(fabs(vx), fmax(fabs(vy), fabs(vz)));
  return fabs(qw) * f;
}

// ---------------------------------------------------------------------------
// FP64 particle
// ---------------------------------------------------------------------------
// Gather, push and position push of particle p.  Fields are read from the
// LDS block when the stencil lies inside it (views fl), else from global
// memory (fg).  Extends the box of the current the particle will deposit.
// an FP64 particle as the push reads it
struct Part64 {
  double x, y, z, ux, uy, uz, gi, w;
};

PXR_DEV Part64 load_part64(const FusedArgs &a, i64 p) {
  return Part64{a.x[p], a.y[p], a.z[p], a.ux[p], a.uy[p], a.uz[p], a.gi[p], a.w[p]};
}

template <int N, bool LO, int PUSHER, int DEPO, typename VL, typename VG>
PXR_DEV void gpd_push(const FusedArgs &a, i64 p, const Part64 &P, const VL *fl, const VG *fg,
                      const int *FR0, const int *fe, Box6 &dep, bool box, double &wabs) {
  constexpr int M = LO ? N - 1 : N;
  const GridGeom &g = a.g;
  const double x = P.x, y = P.y, z = P.z;
  GatherStencil<N, M, double> st;
  gather_stencil<N, M, double>(st, (x - a.xmin) * a.dxi, (y - a.ymin) * a.dyi,
                               (z - a.zmin) * a.dzi, a.s);
  int l2[3], h2[3];
  gather_bounds<N, M, double>(st, l2, h2);
  if (!g.inside(l2[0], l2[1], l2[2]) || !g.inside(h2[0], h2[1], h2[2])) {
    count_bad(a.bad);  // outside the arrays: skipped, as the reference cannot hold it
    return;
  }
  double E[6];
  if (in_block(l2, h2, FR0, fe)) gather_apply<N, M, double>(st, fl, E);
  else gather_apply<N, M, double>(st, fg, E);
  double u[3] = {P.ux, P.uy, P.uz};
  double gi = P.gi;
  if (PUSHER == 1) vay_push<double>(u, gi, E[0], E[1], E[2], E[3], E[4], E[5], a.cst);
  else boris_push<double>(u, gi, E[0], E[1], E[2], E[3], E[4], E[5], a.cst);
  // pxr_pushxyz (boris_3d.F90:880-882)
  const double xn = x + u[0] * gi * a.dt;
  const double yn = y + u[1] * gi * a.dt;
  const double zn = z + u[2] * gi * a.dt;
  a.ux[p] = u[0];
  a.uy[p] = u[1];
  a.uz[p] = u[2];
  a.gi[p] = gi;
  a.x[p] = xn;
  a.y[p] = yn;
  a.z[p] = zn;
  if (!box) return;
  wabs = fmax(wabs, fx_weight<DEPO>(a, a.q * P.w, u[0] * gi, u[1] * gi, u[2] * gi));
  int b0[3], b1[3];
  if (DEPO == 0) {
    EsirkepovIn<double> in;
    esirkepov_prepare<N, double>(in, xn, yn, zn, u[0], u[1], u[2], gi, P.w, a.ec);
    if (!esirkepov_box<N, double>(in, b0, b1)) return;
  } else {
    DirectIn<double> in;
    direct_prepare<N, double>(in, xn, yn, zn, u[0], u[1], u[2], gi, P.w, a.q, a.xmin, a.ymin,
                              a.zmin, a.dt, a.dx, a.dy, a.dz, a.s);
    direct_box<N>(in, b0, b1);
  }
  for (int d = 0; d < 3; d++) {
    dep.lo[d] = min(dep.lo[d], b0[d]);
    dep.hi[d] = max(dep.hi[d], b1[d]);
  }
}

// Extends dep by the box of the current that particle p (in memory as it
// stands) deposits: the deposition-only phase.
template <int N, int DEPO>
PXR_DEV void gpd_box(const FusedArgs &a, i64 p, Box6 &dep, double &wabs) {
  wabs = fmax(wabs, fx_weight<DEPO>(a, a.q * a.w[p], a.ux[p] * a.gi[p], a.uy[p] * a.gi[p],
                                    a.uz[p] * a.gi[p]));
  int b0[3], b1[3];
  if (DEPO == 0) {
    EsirkepovIn<double> in;
    esirkepov_prepare<N, double>(in, a.x[p], a.y[p], a.z[p], a.ux[p], a.uy[p], a.uz[p], a.gi[p],
                                 a.w[p], a.ec);
    if (!esirkepov_box<N, double>(in, b0, b1)) return;
  } else {
    DirectIn<double> in;
    direct_prepare<N, double>(in, a.x[p], a.y[p], a.z[p], a.ux[p], a.uy[p], a.uz[p], a.gi[p],
                              a.w[p], a.q, a.xmin, a.ymin, a.zmin, a.dt, a.dx, a.dy, a.dz, a.s);
    direct_box<N>(in, b0, b1);
  }
  for (int d = 0; d < 3; d++) {
    dep.lo[d] = min(dep.lo[d], b0[d]);
    dep.hi[d] = max(dep.hi[d], b1[d]);
  }
}

// results of a deposit call
constexpr int DEP_OK = 0;   // deposited
constexpr int DEP_BAD = 1;  // the particle left the arrays: skipped (the caller counts it)
constexpr int DEP_OUT = 2;  // its node box is not on the side of the block SEL asks for

// whether node box lo..hi is on the side of block (R0, e) that SEL selects:
// 0 either (no block), 1 inside, 2 not inside
template <int SEL>
PXR_DEV bool sel_block(const int *lo, const int *hi, const int *R0, const int *e) {
  if (SEL == 0) return true;
  const bool in = in_block(lo, hi, R0, e);
  return SEL == 1 ? in : !in;
}

// Deposit pushed particle p (its state is read back from memory).  Returns
// false when the particle left the arrays (counted, skipped).
template <int N, int DEPO, typename Acc>
PXR_DEV bool gpd_deposit(const FusedArgs &a, i64 p, const Acc &acc) {
  const GridGeom &g = a.g;
  int b0[3], b1[3];
  if (DEPO == 0) {
    EsirkepovIn<double> in;
    esirkepov_prepare<N, double>(in, a.x[p], a.y[p], a.z[p], a.ux[p], a.uy[p], a.uz[p], a.gi[p],
                                 a.w[p], a.ec);
    if (!esirkepov_box<N, double>(in, b0, b1) || !g.inside(b0[0], b0[1], b0[2]) ||
        !g.inside(b1[0], b1[1], b1[2]))
      return false;
    esirkepov_deposit<N, double>(in, acc);
  } else {
    DirectIn<double> in;
    direct_prepare<N, double>(in, a.x[p], a.y[p], a.z[p], a.ux[p], a.uy[p], a.uz[p], a.gi[p],
                              a.w[p], a.q, a.xmin, a.ymin, a.zmin, a.dt, a.dx, a.dy, a.dz, a.s);
    direct_box<N>(in, b0, b1);
    if (!g.inside(b0[0], b0[1], b0[2]) || !g.inside(b1[0], b1[1], b1[2])) return false;
    direct_deposit<N, double>(in, acc);
  }
  return true;
}

// ---------------------------------------------------------------------------
// mixed-precision particle: FP64 position, FP32 momentum, gamma and weight
// ---------------------------------------------------------------------------
// node box of the current of a particle at new position pn (cells) with
// velocity v; false if it moved more than one cell (Esirkepov)
template <int N, int DEPO>
PXR_DEV bool mx_box(const FusedArgs &a, const double *pn, const double *v, int *b0, int *b1) {
  const double dts[3] = {a.ec.dtsdx0, a.ec.dtsdy0, a.ec.dtsdz0};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (DEPO == 0) {
      const double uo = pn[d] - dts[d] * v[d];
      const int i0 = eindex<N, double>(pn[d]), i1 = eindex<N, double>(uo);
      const int dd = i1 - i0;
      if (dd < -1 || dd > 1) return false;
      b0[d] = i0 + min(0, dd) + ESup<N>::lo;
      b1[d] = i0 + max(0, dd) + ESup<N>::hi;
    } else {
      const double xm = pn[d] - 0.5 * dts[d] * v[d];
      auto dj = [](double t) { return N == 2 ? f_nint(t) - 1 : f_floor(t) - (N == 1 ? 0 : 1); };
      const int j = dj(xm), j0 = dj(xm - a.s);
      b0[d] = min(j, j0);
      b1[d] = max(j, j0) + N;
    }
  }
  return true;
}

// a mixed-precision particle as the push reads it
struct PartMx {
  double x, y, z;
  float ux, uy, uz, gi, w;
};

PXR_DEV PartMx load_part_mx(const FusedArgs &a, i64 p) {
  return PartMx{a.x[p], a.y[p], a.z[p], a.fux[p], a.fuy[p], a.fuz[p], a.fgi[p], a.fw[p]};
}

template <int N, bool LO, int PUSHER, int DEPO, typename VL, typename VG>
PXR_DEV void gpd_push_mx(const FusedArgs &a, i64 p, const PartMx &P, const VL *fl, const VG *fg,
                         const int *FR0, const int *fe, Box6 &dep, bool box, double &wabs) {
  constexpr int M = LO ? N - 1 : N;
  const GridGeom &g = a.g;
  const double x = P.x, y = P.y, z = P.z;
  GatherStencil<N, M, float> st;
  gather_stencil_mx<N, M>(st, (x - a.xmin) * a.dxi, (y - a.ymin) * a.dyi, (z - a.zmin) * a.dzi,
                          a.s);
  int l2[3], h2[3];
  gather_bounds<N, M, float>(st, l2, h2);
  if (!g.inside(l2[0], l2[1], l2[2]) || !g.inside(h2[0], h2[1], h2[2])) {
    count_bad(a.bad);
    return;
  }
  float E[6];
  if (in_block(l2, h2, FR0, fe)) gather_apply<N, M, float>(st, fl, E);
  else gather_apply<N, M, float>(st, fg, E);
  float u[3] = {P.ux, P.uy, P.uz};
  float gi = P.gi;
  if (PUSHER == 1) vay_push<float>(u, gi, E[0], E[1], E[2], E[3], E[4], E[5], (float)a.cst);
  else boris_push<float>(u, gi, E[0], E[1], E[2], E[3], E[4], E[5], (float)a.cst);
  // position push in FP64 (the products of two floats are exact in FP64)
  const double v[3] = {(double)u[0] * (double)gi, (double)u[1] * (double)gi,
                       (double)u[2] * (double)gi};
  const double xn = x + v[0] * a.dt, yn = y + v[1] * a.dt, zn = z + v[2] * a.dt;
  a.fux[p] = u[0];
  a.fuy[p] = u[1];
  a.fuz[p] = u[2];
  a.fgi[p] = gi;
  a.x[p] = xn;
  a.y[p] = yn;
  a.z[p] = zn;
  if (!box) return;
  wabs = fmax(wabs, fx_weight<DEPO>(a, a.q * (double)P.w, v[0], v[1], v[2]));
  const double pn[3] = {(xn - a.xmin) * a.dxi, (yn - a.ymin) * a.dyi, (zn - a.zmin) * a.dzi};
  int b0[3], b1[3];
  if (!mx_box<N, DEPO>(a, pn, v, b0, b1)) return;
  for (int d = 0; d < 3; d++) {
    dep.lo[d] = min(dep.lo[d], b0[d]);
    dep.hi[d] = max(dep.hi[d], b1[d]);
  }
}

// mixed-precision counterpart of gpd_box
template <int N, int DEPO>
PXR_DEV void gpd_box_mx(const FusedArgs &a, i64 p, Box6 &dep, double &wabs) {
  const float gi = a.fgi[p];
  const double v[3] = {(double)a.fux[p] * (double)gi, (double)a.fuy[p] * (double)gi,
                       (double)a.fuz[p] * (double)gi};
  wabs = fmax(wabs, fx_weight<DEPO>(a, a.q * (double)a.fw[p], v[0], v[1], v[2]));
  const double pn[3] = {(a.x[p] - a.xmin) * a.dxi, (a.y[p] - a.ymin) * a.dyi,
                        (a.z[p] - a.zmin) * a.dzi};
  int b0[3], b1[3];
  if (!mx_box<N, DEPO>(a, pn, v, b0, b1)) return;
  for (int d = 0; d < 3; d++) {
    dep.lo[d] = min(dep.lo[d], b0[d]);
    dep.hi[d] = max(dep.hi[d], b1[d]);
  }
}

template <int N, int DEPO, int SEL = 0, typename Acc>
PXR_DEV int gpd_deposit_sel_mx(const FusedArgs &a, i64 p, const Acc &acc, const int *R0,
                               const int *e) {
  const GridGeom &g = a.g;
  const float gi = a.fgi[p];
  const float uf[3] = {a.fux[p], a.fuy[p], a.fuz[p]};
  const double v[3] = {(double)uf[0] * (double)gi, (double)uf[1] * (double)gi,
                       (double)uf[2] * (double)gi};
  const double pn[3] = {(a.x[p] - a.xmin) * a.dxi, (a.y[p] - a.ymin) * a.dyi,
                        (a.z[p] - a.zmin) * a.dzi};
  int b0[3], b1[3];
  if (!mx_box<N, DEPO>(a, pn, v, b0, b1) || !g.inside(b0[0], b0[1], b0[2]) ||
      !g.inside(b1[0], b1[1], b1[2]))
    return DEP_BAD;
  if (!sel_block<SEL>(b0, b1, R0, e)) return DEP_OUT;
  const float wq = (float)(a.q * (double)a.fw[p]);
  if (DEPO == 0) {
    EAxisMx<N> X, Y, Z;
    esirkepov_axis_mx<N>(pn[0], pn[0] - a.ec.dtsdx0 * v[0], X);
    esirkepov_axis_mx<N>(pn[1], pn[1] - a.ec.dtsdy0 * v[1], Y);
    esirkepov_axis_mx<N>(pn[2], pn[2] - a.ec.dtsdz0 * v[2], Z);
    esirkepov_accum_mx<N>(X, Y, Z, wq * (float)a.ec.invdtdx, wq * (float)a.ec.invdtdy,
                          wq * (float)a.ec.invdtdz, acc);
  } else {
    DirectIn<float> in;
    const float vf[3] = {uf[0] * gi, uf[1] * gi, uf[2] * gi};
    const double h[3] = {0.5 * a.ec.dtsdx0, 0.5 * a.ec.dtsdy0, 0.5 * a.ec.dtsdz0};
    const double invvol = a.dxi * a.dyi * a.dzi;
    direct_prepare_mx<N>(in, pn, vf, h, (float)(a.q * (double)a.fw[p] * invvol), a.s);
    direct_deposit<N, float>(in, acc);
  }
  return DEP_OK;
}

// Deposit pushed mixed-precision particle p.  Returns false when the particle
// left the arrays (counted, skipped).
template <int N, int DEPO, typename Acc>
PXR_DEV bool gpd_deposit_mx(const FusedArgs &a, i64 p, const Acc &acc) {
  return gpd_deposit_sel_mx<N, DEPO>(a, p, acc, nullptr, nullptr) == DEP_OK;
}

// Fixed-point scale of a work item of n particles whose largest fx_weight
// has the bits wbits: a power of two under which every contribution is
// below 2^50 in magnitude and a node's sum of at most n of them stays below
// 2^62 (the bound is a.fxb times that weight, so the quantum follows the
// size of the item's actual contributions).
PXR_DEV void fx_scale(const FusedArgs &a, unsigned long long wbits, int n, double &fxs,
                      double &fxs_inv) {
  fxs = fxs_inv = 1.0;
  const double bound = a.fxb * __longlong_as_double((long long)wbits);
  if (bound > 0.0) {
    int eb;
    frexp(bound, &eb);  // bound < 2^eb
    const int en = n > 1 ? 32 - __clz(n - 1) : 0;  // n <= 2^en
    const int k = min(50 - eb, 62 - eb - en);
    fxs = ldexp(1.0, k);
    fxs_inv = ldexp(1.0, -k);
  }
}

// ---------------------------------------------------------------------------
// the kernels
// ---------------------------------------------------------------------------
// Persistent kernel over the work items.  Per work item:
//   1. the node box of its field block, from the sorting-grid cells that
//      hold its particles (or, without sorting information, from one pass
//      over the positions); the six field blocks are staged into LDS (FP64,
//      or FP32 for a mixed-precision species);
//   2. gather (from LDS; a particle whose stencil leaves the block, e.g.
//      after a push without re-sorting, reads global memory), push and
//      position push, lanes on consecutive particles; the box of the
//      currents is reduced across the workgroup;
//   3. the current block is zeroed in LDS and the pushed particles are
//      deposited with LDS atomics; lanes of a wave take particles S apart
//      (one or more cells apart in a cell-sorted run) so that their atomics
//      rarely meet;
//   4. the current block is added to global J (FP64 atomics).
// A work item whose field block does not fit in LDS is left whole to
// k_fused_overflow; one whose current block does not fit is left to it
// after the push.
// a.phase selects the phases: PHASE_PUSH alone is the gather and push of
// field_gathering_plus_particle_pusher; PHASE_DEPOSIT alone deposits the
// particles as they stand (pxrdepose_currents_on_grid_jxjyjz after
// particle_bcs), the box of step 3 then coming from one pass over them.
// Registers are capped for three waves per SIMD (three workgroups of 256
// lanes per CU), except for the order-3 direct deposition of an FP64
// species, which would spill and keeps two.
template <int BS, int N, bool LO, int PUSHER, int DEPO, bool MX, int JACC>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(
    (N == 3 && DEPO == 3 && !MX) ? 2 : 3))) void k_fused(FusedArgs a) {
  typedef typename std::conditional<MX, float, double>::type FT;  // field block
  // current block: FP64, 64-bit fixed point (JACC 3), or FP32 for the
  // mixed-precision modes 0 and 1
  typedef typename std::conditional<
      JACC == 3, unsigned long long,
      typename std::conditional<MX && JACC != 2, float, double>::type>::type JT;
  extern __shared__ double smem[];
  __shared__ int s_chunk;
  __shared__ int s_box[6];
  __shared__ unsigned long long s_wabs;  // largest |q w| of the work item (JACC 3)
  const GridGeom &g = a.g;
  const int glo[3] = {-g.gx, -g.gy, -g.gz};
  const int ghi[3] = {g.nx + g.gx, g.ny + g.gy, g.nz + g.gz};
  const double x0[3] = {a.xmin, a.ymin, a.zmin}, di[3] = {a.dxi, a.dyi, a.dzi};
  const int nch = *a.nchunks;
  for (;;) {
    if (threadIdx.x == 0) {
      const int q = atomicAdd(a.queue, 1);
      s_chunk = (a.order && q < nch) ? a.order[q] : q;
      for (int d = 0; d < 3; d++) {
        s_box[d] = INT_MAX;
        s_box[3 + d] = INT_MIN;
      }
      s_wabs = 0ull;
    }
    __syncthreads();
    const int c = s_chunk;
    if (c >= nch) break;
    const Chunk ch = a.chunks[c];
    const i64 p0 = ch.begin;
    const bool push = a.phase & PHASE_PUSH, depo = a.phase & PHASE_DEPOSIT;
    Box6 dep;
    box_reset(dep);
    double wabs = 0.0;
    if (push) {
      // ---- 1. field block -----------------------------------------------
      double ulo[3], uhi[3];
      if (ch.z >= 0) {
        rows_range(a.bg, ch.tile, ch.z, ch.z1, ch.y0, ch.y1, ulo, uhi);
      } else {
        // no sorting information: one pass over the positions, in whole cells
        Box6 b;
        box_reset(b);
        for (int k = threadIdx.x; k < ch.n; k += BS) {
          const i64 p = p0 + k;
          const double pp[3] = {a.x[p], a.y[p], a.z[p]};
          for (int d = 0; d < 3; d++) {
            const int cc = f_floor((pp[d] - x0[d]) * di[d]);
            b.lo[d] = min(b.lo[d], cc);
            b.hi[d] = max(b.hi[d], cc);
          }
        }
        wave_minmax(b);
        if ((threadIdx.x & 63) == 0)
          for (int d = 0; d < 3; d++) {
            atomicMin(&s_box[d], b.lo[d]);
            atomicMax(&s_box[3 + d], b.hi[d]);
          }
        __syncthreads();
        for (int d = 0; d < 3; d++) {
          const int lo = max(s_box[d], glo[d] - 4), hi = min(s_box[3 + d], ghi[d] + 4);
          ulo[d] = lo - 1e-7;
          uhi[d] = hi + 1.0 + 1e-7;
        }
        __syncthreads();
        if (threadIdx.x == 0)
          for (int d = 0; d < 3; d++) {
            s_box[d] = INT_MAX;
            s_box[3 + d] = INT_MIN;
          }
      }
      int FR0[3], fe[3];
      for (int d = 0; d < 3; d++) {
        int lo, hi;
        gather_range(a.bg, ulo[d], uhi[d], lo, hi);
        FR0[d] = max(lo, glo[d]);
        fe[d] = max(min(hi, ghi[d]) - FR0[d] + 1, 0);
      }
      const i64 nb64 = (i64)fe[0] * fe[1] * fe[2];
      if (nb64 <= 0 || 6 * nb64 * (i64)sizeof(FT) > a.lds_bytes) {
        if (threadIdx.x == 0) {
          const int slot = atomicAdd(a.ovf_count, 1);
          if (slot < a.ovf_cap) a.ovf[slot] = make_int2(c, 0);
        }
        __syncthreads();
        continue;
      }
      const int nb = (int)nb64;
      FT *sf = reinterpret_cast<FT *>(smem);
      // one node per lane and step, its six field values loaded together:
      // the node's position is decomposed once for all six, and the six
      // loads are in flight at the same time (two nodes per step, twelve
      // loads, made the launch slower)
      for (int idx = threadIdx.x; idx < nb; idx += BS) {
        const int i = FR0[0] + idx % fe[0];
        const int j = FR0[1] + (idx / fe[0]) % fe[1];
        const int k = FR0[2] + idx / (fe[0] * fe[1]);
        const i64 go = g.idx(i, j, k);
        double v[6];
#pragma unroll
        for (int f = 0; f < 6; f++) v[f] = a.F[f][go];
#pragma unroll
        for (int f = 0; f < 6; f++) sf[f * nb + idx] = (FT)v[f];
      }
      __syncthreads();
      // ---- 2. gather, push, position push -------------------------------
      LdsView<FT> fl[6];
      for (int f = 0; f < 6; f++)
        fl[f] = LdsView<FT>{sf + f * nb, FR0[0], FR0[1], FR0[2], fe[0], fe[0] * fe[1]};
      if constexpr (MX) {
        FieldViewF fg[6];
        for (int f = 0; f < 6; f++) fg[f] = FieldViewF{a.F[f] + g.idx(0, 0, 0), g.NX, g.NX * g.NY};
        // the next particle's values are loaded before the current one is
        // pushed, so that their round trip overlaps its arithmetic
        int k = threadIdx.x;
        if (k < ch.n) {
          PartMx cur = load_part_mx(a, p0 + k);
          for (;;) {
            const int kn = k + BS;
            PartMx nxt = cur;
            if (kn < ch.n) nxt = load_part_mx(a, p0 + kn);
            gpd_push_mx<N, LO, PUSHER, DEPO>(a, p0 + k, cur, fl, fg, FR0, fe, dep, depo, wabs);
            if (kn >= ch.n) break;
            k = kn;
            cur = nxt;
          }
        }
      } else {
        FieldView<double> fg[6];
        for (int f = 0; f < 6; f++)
          fg[f] = FieldView<double>{a.F[f] + g.idx(0, 0, 0), g.NX, g.NX * g.NY};
        int k = threadIdx.x;
        if (k < ch.n) {
          Part64 cur = load_part64(a, p0 + k);
          for (;;) {
            const int kn = k + BS;
            Part64 nxt = cur;
            if (kn < ch.n) nxt = load_part64(a, p0 + kn);
            gpd_push<N, LO, PUSHER, DEPO>(a, p0 + k, cur, fl, fg, FR0, fe, dep, depo, wabs);
            if (kn >= ch.n) break;
            k = kn;
            cur = nxt;
          }
        }
      }
      if (!depo) {
        __syncthreads();  // field block no longer read
        continue;
      }
    } else {
      // deposition only: the box of the currents of the particles as they are
      for (int k = threadIdx.x; k < ch.n; k += BS) {
        if constexpr (MX) gpd_box_mx<N, DEPO>(a, p0 + k, dep, wabs);
        else gpd_box<N, DEPO>(a, p0 + k, dep, wabs);
      }
    }
    wave_minmax(dep);
    if constexpr (JACC == 3) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) wabs = fmax(wabs, __shfl_xor(wabs, o));
    }
    __syncthreads();  // field block no longer read
    if ((threadIdx.x & 63) == 0) {
      for (int d = 0; d < 3; d++) {
        atomicMin(&s_box[d], dep.lo[d]);
        atomicMax(&s_box[3 + d], dep.hi[d]);
      }
      // non-negative doubles order as their bit patterns do
      if (JACC == 3) atomicMax(&s_wabs, (unsigned long long)__double_as_longlong(wabs));
    }
    __syncthreads();
    // fixed point (JACC 3): the scale of the work item
    double fxs = 1.0, fxs_inv = 1.0;
    if constexpr (JACC == 3) fx_scale(a, s_wabs, ch.n, fxs, fxs_inv);
    int R0[3], e[3];
    for (int d = 0; d < 3; d++) {
      R0[d] = max(s_box[d], glo[d]);
      const int R1 = min(s_box[3 + d], ghi[d]);
      e[d] = R1 >= R0[d] ? R1 - R0[d] + 1 : 0;
    }
    // LDS layout of the current block (optionally odd row and plane pitch)
    const int py = a.jpad ? (e[0] | 1) : e[0];
    const int pzr = py * e[1];
    const int pz = a.jpad ? (pzr | 1) : pzr;
    const i64 nj64 = (i64)pz * e[2];
    const int narr = (MX && JACC == 1) ? 6 : 3;  // with the compensation arrays
    const i64 ne = (i64)e[0] * e[1] * e[2];
    if (ne <= 0 || narr * nj64 * (i64)sizeof(JT) > a.j_bytes) {
      // pushed; the deposition is left to k_fused_cut (stage 3: a block cut
      // down to fit) or to the overflow kernel (stage 1: global memory)
      if (threadIdx.x == 0 && ne > 0) {
        const bool cut = JACC == 3 && MX && (a.phase & PHASE_CUT) && ch.z >= 0;
        const int slot = atomicAdd(a.ovf_count, 1);
        if (slot < a.ovf_cap) a.ovf[slot] = make_int2(c, cut ? 3 : 1);
      }
      __syncthreads();
      continue;
    }
    const int nj = (int)nj64;
    JT *sj = reinterpret_cast<JT *>(smem);
    for (int idx = threadIdx.x; idx < narr * nj; idx += BS) sj[idx] = (JT)0;
    __syncthreads();
    // ---- 3. deposit -------------------------------------------------------
    {
      // lanes take particles S apart inside super-blocks of 64*S particles
      int S = (ch.n + 63) >> 6;
      const int cap = a.sstride > 0 ? a.sstride : ch.s;
      if (cap > 0 && cap < S) S = cap;
      const int SB = 64 * S;
      const int nkk = ((ch.n + SB - 1) / SB) * SB;
      auto pick = [&](int kk) {
        const int r = kk % SB;
        return a.spread ? (kk - r) + (r & 63) * S + (r >> 6) : kk;
      };
      if constexpr (MX) {
        float *se = reinterpret_cast<float *>(smem) + 3 * nj;  // compensation (JACC 1)
        const LdsAccMx<JACC> la{{sj, sj + nj, sj + 2 * nj}, {se, se + nj, se + 2 * nj},
                                R0[0], R0[1], R0[2], py, pz, fxs};
        // (loading the next particle ahead, as the push loop does, makes this
        // loop spill registers and slows the launch)
        for (int kk = threadIdx.x; kk < nkk; kk += BS) {
          const int k = pick(kk);
          if (k < ch.n && !gpd_deposit_mx<N, DEPO>(a, p0 + k, la)) count_bad(a.bad);
        }
      } else if constexpr (JACC == 3) {
        const LdsAccFx la{{sj, sj + nj, sj + 2 * nj}, R0[0], R0[1], R0[2], py, pz, fxs};
        for (int kk = threadIdx.x; kk < nkk; kk += BS) {
          const int k = pick(kk);
          if (k < ch.n && !gpd_deposit<N, DEPO>(a, p0 + k, la)) count_bad(a.bad);
        }
      } else {
        double *sjd = reinterpret_cast<double *>(smem);
        const LdsAcc la{{sjd, sjd + nj, sjd + 2 * nj}, R0[0], R0[1], R0[2], py, pz};
        for (int kk = threadIdx.x; kk < nkk; kk += BS) {
          const int k = pick(kk);
          if (k < ch.n && !gpd_deposit<N, DEPO>(a, p0 + k, la)) count_bad(a.bad);
        }
      }
    }
    __syncthreads();
    // ---- 4. flush ----------------------------------------------------------
    for (int idx = threadIdx.x; idx < (int)ne; idx += BS) {
      const int i = idx % e[0];
      const int j = (idx / e[0]) % e[1];
      const int k = idx / (e[0] * e[1]);
      const i64 gi_ = g.idx(R0[0] + i, R0[1] + j, R0[2] + k);
      const int li = i + py * j + pz * k;
      for (int c3 = 0; c3 < 3; c3++) {
        double v;
        if constexpr (JACC == 3) v = (double)(long long)sj[c3 * nj + li] * fxs_inv;
        else v = (double)sj[c3 * nj + li];
        if (MX && JACC == 1) v += (double)sj[(3 + c3) * nj + li];
        if (v != 0.0) unsafeAtomicAdd(a.J[c3] + gi_, v);
      }
    }
    __syncthreads();
  }
}

// Work items whose blocks did not fit in LDS: gather from global memory and
// deposit with global atomics, one workgroup per work item.  Stage 1: the
// particles were already pushed by k_fused.
template <int BS, int N, bool LO, int PUSHER, int DEPO, bool MX>
__global__ __launch_bounds__(BS) void k_fused_overflow(FusedArgs a) {
  const int n = min(*a.ovf_count, a.ovf_cap);
  const GridGeom &g = a.g;
  const int none[3] = {0, 0, 0};  // empty block: every gather reads global memory
  for (int o = blockIdx.x; o < n; o += gridDim.x) {
    const int2 ov = a.ovf[o];
    if (ov.y == 3) continue;  // k_fused_cut
    const Chunk ch = a.chunks[ov.x];
    if (ov.y == 0 && (a.phase & PHASE_PUSH)) {
      Box6 dep;
      box_reset(dep);
      double wabs = 0.0;
      if constexpr (MX) {
        FieldViewF fg[6];
        for (int f = 0; f < 6; f++) fg[f] = FieldViewF{a.F[f] + g.idx(0, 0, 0), g.NX, g.NX * g.NY};
        for (int k = threadIdx.x; k < ch.n; k += BS)
          gpd_push_mx<N, LO, PUSHER, DEPO>(a, ch.begin + k, load_part_mx(a, ch.begin + k), fg, fg,
                                           none, none, dep, false, wabs);
      } else {
        FieldView<double> fg[6];
        for (int f = 0; f < 6; f++)
          fg[f] = FieldView<double>{a.F[f] + g.idx(0, 0, 0), g.NX, g.NX * g.NY};
        for (int k = threadIdx.x; k < ch.n; k += BS)
          gpd_push<N, LO, PUSHER, DEPO>(a, ch.begin + k, load_part64(a, ch.begin + k), fg, fg, none,
                                        none, dep, false, wabs);
      }
    }
    __syncthreads();
    if (!(a.phase & PHASE_DEPOSIT)) continue;
    const int S = (ch.n + 63) >> 6;
    for (int kk = threadIdx.x; kk < S * 64; kk += BS) {
      const int k = a.spread ? (kk & 63) * S + (kk >> 6) : kk;
      if (k >= ch.n) continue;
      bool ok;
      if constexpr (MX) ok = gpd_deposit_mx<N, DEPO>(a, ch.begin + k, GAccF{{a.J[0], a.J[1], a.J[2]}, g});
      else ok = gpd_deposit<N, DEPO>(a, ch.begin + k, GAcc{{a.J[0], a.J[1], a.J[2]}, g});
      if (!ok) count_bad(a.bad);
    }
    __syncthreads();
  }
}

// Work items of a mixed-precision species whose current box did not fit in
// LDS (stage 3: pushed by k_fused).  The box of the pushed particles and
// the fixed-point scale are found again as k_fused found them, in a pass
// over the particles.  The current block is cut down to the block of the
// particles' old positions, as the work-item builder sized it, grown
// towards that box while it fits; the particles whose box leaves it add
// into global memory.  One workgroup per work item.
template <int BS, int N, bool LO, int PUSHER, int DEPO>
__global__ __launch_bounds__(BS) void k_fused_cut(FusedArgs a) {
  extern __shared__ double smem[];
  __shared__ int s_box[6];
  __shared__ unsigned long long s_wabs;
  typedef unsigned long long JT;
  const int n = min(*a.ovf_count, a.ovf_cap);
  const GridGeom &g = a.g;
  const int glo[3] = {-g.gx, -g.gy, -g.gz};
  const int ghi[3] = {g.nx + g.gx, g.ny + g.gy, g.nz + g.gz};
  auto layout = [&](const int *ee, int &py_, int &pz_) -> i64 {
    py_ = a.jpad ? (ee[0] | 1) : ee[0];
    const int pzr = py_ * ee[1];
    pz_ = a.jpad ? (pzr | 1) : pzr;
    return (i64)pz_ * ee[2];
  };
  for (int o = blockIdx.x; o < n; o += gridDim.x) {
    const int2 ov = a.ovf[o];
    if (ov.y != 3) continue;
    const Chunk ch = a.chunks[ov.x];
    // the box of the currents and the largest fixed-point weight
    if (threadIdx.x == 0) {
      for (int d = 0; d < 3; d++) {
        s_box[d] = INT_MAX;
        s_box[3 + d] = INT_MIN;
      }
      s_wabs = 0ull;
    }
    __syncthreads();
    {
      Box6 dep;
      box_reset(dep);
      double wabs = 0.0;
      for (int k = threadIdx.x; k < ch.n; k += BS) gpd_box_mx<N, DEPO>(a, ch.begin + k, dep, wabs);
      wave_minmax(dep);
#pragma unroll
      for (int sh = 32; sh > 0; sh >>= 1) wabs = fmax(wabs, __shfl_xor(wabs, sh));
      if ((threadIdx.x & 63) == 0) {
        for (int d = 0; d < 3; d++) {
          atomicMin(&s_box[d], dep.lo[d]);
          atomicMax(&s_box[3 + d], dep.hi[d]);
        }
        atomicMax(&s_wabs, (unsigned long long)__double_as_longlong(wabs));
      }
    }
    __syncthreads();
    int B0[3], B1[3];  // the box, clamped to the arrays (empty when B1 < B0)
    for (int d = 0; d < 3; d++) {
      B0[d] = max(s_box[d], glo[d]);
      B1[d] = min(s_box[3 + d], ghi[d]);
    }
    const unsigned long long wbits = s_wabs;
    __syncthreads();  // s_box and s_wabs are read before the next work item resets them
    // the old positions' block, intersected with the box of the particles
    double ulo[3], uhi[3];
    rows_range(a.bg, ch.tile, ch.z, ch.z1, ch.y0, ch.y1, ulo, uhi);
    int lo[3], hi[3], ee[3];
    bool any = true;
    for (int d = 0; d < 3; d++) {
      int l, h;
      deposit_range(a.bg, ulo[d], uhi[d], a.bg.mv[d], l, h);
      lo[d] = max(max(l, glo[d]), B0[d]);
      hi[d] = min(min(h, ghi[d]), B1[d]);
      any = any && hi[d] >= lo[d];
    }
    auto fits = [&]() {
      int q0, q1;
      for (int d = 0; d < 3; d++) ee[d] = hi[d] - lo[d] + 1;
      return 3 * layout(ee, q0, q1) * (i64)sizeof(JT) <= a.j_bytes;
    };
    const GAccF ga{{a.J[0], a.J[1], a.J[2]}, g};
    if (!(any && fits())) {
      // no block fits: every particle into global memory
      for (int k = threadIdx.x; k < ch.n; k += BS)
        if (!gpd_deposit_mx<N, DEPO>(a, ch.begin + k, ga)) count_bad(a.bad);
      continue;
    }
    for (int it = 0; it < 4; it++) {
      bool grown = false;
      for (int d = 0; d < 3; d++) {
        if (lo[d] > B0[d]) {
          lo[d]--;
          if (fits()) grown = true;
          else lo[d]++;
        }
        if (hi[d] < B1[d]) {
          hi[d]++;
          if (fits()) grown = true;
          else hi[d]--;
        }
      }
      if (!grown) break;
    }
    int R0[3], e[3];
    for (int d = 0; d < 3; d++) {
      R0[d] = lo[d];
      e[d] = hi[d] - lo[d] + 1;
    }
    int py, pz;
    const int nj = (int)layout(e, py, pz);
    double fxs, fxs_inv;
    fx_scale(a, wbits, ch.n, fxs, fxs_inv);
    JT *sj = reinterpret_cast<JT *>(smem);
    for (int idx = threadIdx.x; idx < 3 * nj; idx += BS) sj[idx] = 0ull;
    __syncthreads();
    const LdsAccMx<3> la{{sj, sj + nj, sj + 2 * nj}, {nullptr, nullptr, nullptr},
                         R0[0], R0[1], R0[2], py, pz, fxs};
    for (int k = threadIdx.x; k < ch.n; k += BS) {
      const i64 p = ch.begin + k;
      const int r = gpd_deposit_sel_mx<N, DEPO, 1>(a, p, la, R0, e);
      if (r == DEP_BAD) count_bad(a.bad);
      else if (r == DEP_OUT) gpd_deposit_mx<N, DEPO>(a, p, ga);
    }
    __syncthreads();
    const int ne = e[0] * e[1] * e[2];
    for (int idx = threadIdx.x; idx < ne; idx += BS) {
      const int i = idx % e[0];
      const int j = (idx / e[0]) % e[1];
      const int k = idx / (e[0] * e[1]);
      const i64 gi_ = g.idx(R0[0] + i, R0[1] + j, R0[2] + k);
      const int li = i + py * j + pz * k;
      for (int c3 = 0; c3 < 3; c3++) {
        const double v = (double)(long long)sj[c3 * nj + li] * fxs_inv;
        if (v != 0.0) unsafeAtomicAdd(a.J[c3] + gi_, v);
      }
    }
    __syncthreads();
  }
}

// launches k_fused and k_fused_overflow for the combination `key`
// (order*100 + lower-order-in-v*10 + pusher*2 + direct); defined by the
// FP64 and the mixed-precision translation units
int launch_fused_fp64(int key, const FusedArgs &a, int nwg, size_t lds, hipStream_t st);
int launch_fused_mixed_f32(int key, const FusedArgs &a, int nwg, size_t lds, hipStream_t st);
int launch_fused_mixed_f64acc(int key, const FusedArgs &a, int nwg, size_t lds, hipStream_t st);
int launch_fused_fp64_fx(int key, const FusedArgs &a, int nwg, size_t lds, hipStream_t st);
int launch_fused_mixed_fx(int key, const FusedArgs &a, int nwg, size_t lds, hipStream_t st);

}  // namespace pxr
