// pxr_kernels.h -- per-particle arithmetic of the hot path, shared by the
// leaf kernels (global memory) and the tiled fused kernel (LDS blocks).
//
// Gather:  field_gathering_o{1,2,3}_3d.F90 (scalar routines) -- summed in
//          factored form, x fastest, then y, then z.
// Pushers: boris_3d.F90:51-141, 1059-1107; vay_3d.F90:56-241 -- the
//          reference's operation order, no contraction.
// Esirkepov: esirkepov_3d.F90:82/1734/3263 -- shapes and cell indices in the
//          reference's operation order; the running sums along each axis are
//          formed as (prefix sum of dS) x (transverse weight), which equals
//          the reference's running sum of products up to rounding.
// Direct:  direct_current_deposition_3d.F90:99/879/1992.
#pragma once
#include "pxr_device.h"

namespace pxr {

template <typename T> struct FieldView {
  const T *p;  // points at node (0,0,0)
  i64 sy, sz;
  PXR_DEV T at(int i, int j, int k) const { return p[i + sy * j + sz * k]; }
};

// LDS block view: element (i,j,k) of the block whose first node is
// (i0,j0,k0) with row length sy and plane size sz.
template <typename T> struct LdsView {
  const T *p;
  int i0, j0, k0, sy, sz;
  PXR_DEV T at(int i, int j, int k) const {
    return p[(i - i0) + sy * (j - j0) + sz * (k - k0)];
  }
};

// ---------------------------------------------------------------------------
// gather
// ---------------------------------------------------------------------------
template <int N, int M, typename T> struct GatherStencil {
  int j, k, l, j0, k0, l0;
  T sx[N + 1], sy[N + 1], sz[N + 1];
  T sx0[M + 1], sy0[M + 1], sz0[M + 1];
};

template <int N, int M, typename T>
PXR_DEV void gather_stencil(GatherStencil<N, M, T> &st, T x, T y, T z, T s) {
  st.j = shape_nodal<N, T>(x, st.sx);
  st.k = shape_nodal<N, T>(y, st.sy);
  st.l = shape_nodal<N, T>(z, st.sz);
  st.j0 = shape_stag<M, T>(x, s, st.sx0);
  st.k0 = shape_stag<M, T>(y, s, st.sy0);
  st.l0 = shape_stag<M, T>(z, s, st.sz0);
}

template <int N, int M, typename T>
PXR_DEV void gather_bounds(const GatherStencil<N, M, T> &st, int *lo, int *hi) {
  lo[0] = min(st.j, st.j0);
  lo[1] = min(st.k, st.k0);
  lo[2] = min(st.l, st.l0);
  hi[0] = max(st.j + N, st.j0 + M);
  hi[1] = max(st.k + N, st.k0 + M);
  hi[2] = max(st.l + N, st.l0 + M);
}

// sum_c wc[c] sum_b wb[b] sum_a wa[a] f(i0+a, j0+b, k0+c)
template <int NA, int NB, int NC, typename T, typename V>
PXR_DEV T gsum(const V &f, int i0, int j0, int k0, const T *wa, const T *wb, const T *wc) {
  T acc = T(0);
#pragma unroll
  for (int c = 0; c <= NC; c++) {
    T ab = T(0);
#pragma unroll
    for (int b = 0; b <= NB; b++) {
      T aa = T(0);
#pragma unroll
      for (int a = 0; a <= NA; a++) aa = fma(wa[a], f.at(i0 + a, j0 + b, k0 + c), aa);
      ab = fma(wb[b], aa, ab);
    }
    acc = fma(wc[c], ab, acc);
  }
  return acc;
}

template <int N, int M, typename T, typename V>
PXR_DEV void gather_apply(const GatherStencil<N, M, T> &st, const V *f, T *out) {
  out[0] = gsum<M, N, N, T>(f[0], st.j0, st.k, st.l, st.sx0, st.sy, st.sz);    // Ex
  out[1] = gsum<N, M, N, T>(f[1], st.j, st.k0, st.l, st.sx, st.sy0, st.sz);    // Ey
  out[2] = gsum<N, N, M, T>(f[2], st.j, st.k, st.l0, st.sx, st.sy, st.sz0);    // Ez
  out[3] = gsum<N, M, M, T>(f[3], st.j, st.k0, st.l0, st.sx, st.sy0, st.sz0);  // Bx
  out[4] = gsum<M, N, M, T>(f[4], st.j0, st.k, st.l0, st.sx0, st.sy, st.sz0);  // By
  out[5] = gsum<M, M, N, T>(f[5], st.j0, st.k0, st.l, st.sx0, st.sy0, st.sz);  // Bz
}

// global-memory gather with a bounds check against the array geometry
template <int N, int M, typename T>
PXR_DEV bool gather_eb(T x, T y, T z, T s, const GridGeom &g, const FieldView<T> *v, T *out) {
  GatherStencil<N, M, T> st;
  gather_stencil<N, M, T>(st, x, y, z, s);
  int lo[3], hi[3];
  gather_bounds<N, M, T>(st, lo, hi);
  if (!g.inside(lo[0], lo[1], lo[2]) || !g.inside(hi[0], hi[1], hi[2])) return false;
  gather_apply<N, M, T>(st, v, out);
  return true;
}

// ---------------------------------------------------------------------------
// pushers (clight = 1, normalised units)
// ---------------------------------------------------------------------------
template <typename T> PXR_DEV T gamma_inv(T ux, T uy, T uz) {
  T usq = (ux * ux + uy * uy + uz * uz);
  return T(1) / sqrt(T(1) + usq);
}

// boris_3d.F90:95-131; cst = q*dt*0.5/m
template <typename T>
PXR_DEV void boris_push(T *u, T &gi, T ex, T ey, T ez, T bx, T by, T bz, T cst) {
  T ux = u[0] + ex * cst, uy = u[1] + ey * cst, uz = u[2] + ez * cst;
  T usq = (ux * ux + uy * uy + uz * uz);
  T g = T(1) / sqrt(T(1) + usq);
  T tx = g * bx * cst, ty = g * by * cst, tz = g * bz * cst;
  T tsqi = T(2) / (T(1) + tx * tx + ty * ty + tz * tz);
  T sx = tx * tsqi, sy = ty * tsqi, sz = tz * tsqi;
  T uxp = ux + uy * tz - uz * ty;
  T uyp = uy + uz * tx - ux * tz;
  T uzp = uz + ux * ty - uy * tx;
  ux = ux + uyp * sz - uzp * sy;
  uy = uy + uzp * sx - uxp * sz;
  uz = uz + uxp * sy - uyp * sx;
  ux = ux + ex * cst;
  uy = uy + ey * cst;
  uz = uz + ez * cst;
  usq = (ux * ux + uy * uy + uz * uz);
  gi = T(1) / sqrt(T(1) + usq);
  u[0] = ux;
  u[1] = uy;
  u[2] = uz;
}

// common tail of the Vay push (vay_3d.F90:113-131)
template <typename T>
PXR_DEV void vay_tail(T *u, T &gi, T uxpr, T uypr, T uzpr, T bx, T by, T bz, T bconst) {
  T taux = bconst * bx, tauy = bconst * by, tauz = bconst * bz;
  T tausq = taux * taux + tauy * tauy + tauz * tauz;
  T gprsq = (T(1) + (uxpr * uxpr + uypr * uypr + uzpr * uzpr));
  T ust = (uxpr * taux + uypr * tauy + uzpr * tauz);
  T sigma = gprsq - tausq;
  T gisq = T(2) / (sigma + sqrt(sigma * sigma + T(4) * (tausq + ust * ust)));
  T g = sqrt(gisq);
  gi = g;
  T bg = bconst * g;
  T tx = bg * bx, ty = bg * by, tz = bg * bz;
  T s = T(1) / (T(1) + tausq * gisq);
  T tu = tx * uxpr + ty * uypr + tz * uzpr;
  u[0] = s * (uxpr + tx * tu + uypr * tz - uzpr * ty);
  u[1] = s * (uypr + ty * tu + uzpr * tx - uxpr * tz);
  u[2] = s * (uzpr + tz * tu + uxpr * ty - uypr * tx);
}

// which = 0 (vay_3d.F90:86-131): cst = q*dt/m, bconst = cst/2, old gi read
template <typename T>
PXR_DEV void vay_push(T *u, T &gi, T ex, T ey, T ez, T bx, T by, T bz, T cst) {
  const T bconst = T(0.5) * cst;
  T taux = bconst * bx, tauy = bconst * by, tauz = bconst * bz;
  T g = gi;
  T uxpr = u[0] + cst * ex + (u[1] * tauz - u[2] * tauy) * g;
  T uypr = u[1] + cst * ey + (u[2] * taux - u[0] * tauz) * g;
  T uzpr = u[2] + cst * ez + (u[0] * tauy - u[1] * taux) * g;
  vay_tail<T>(u, gi, uxpr, uypr, uzpr, bx, by, bz, bconst);
}

// which = 1 (vay_3d.F90:139-160): cst = q*dt/(2m)
template <typename T>
PXR_DEV void vay_half1(T *u, T &gi, T ex, T ey, T ez, T bx, T by, T bz, T cst) {
  T vx = u[0] * gi, vy = u[1] * gi, vz = u[2] * gi;
  T ux = u[0] + cst * (ex + vy * bz - vz * by);
  T uy = u[1] + cst * (ey + vz * bx - vx * bz);
  T uz = u[2] + cst * (ez + vx * by - vy * bx);
  gi = T(1) / sqrt(T(1) + (ux * ux + uy * uy + uz * uz));
  u[0] = ux;
  u[1] = uy;
  u[2] = uz;
}

// which = 2 (vay_3d.F90:166-204): cst = bconst = q*dt/(2m)
template <typename T>
PXR_DEV void vay_half2(T *u, T &gi, T ex, T ey, T ez, T bx, T by, T bz, T cst) {
  T uxpr = u[0] + cst * ex, uypr = u[1] + cst * ey, uzpr = u[2] + cst * ez;
  vay_tail<T>(u, gi, uxpr, uypr, uzpr, bx, by, bz, cst);
}

// ---------------------------------------------------------------------------
// Esirkepov deposition
// ---------------------------------------------------------------------------
template <typename T> struct EsirkConst {
  T xmin, ymin, zmin, dxi, dyi, dzi;
  T dtsdx0, dtsdy0, dtsdz0;     // dt/dx ...
  T invdtdx, invdtdy, invdtdz;  // 1/(dt*dy*dz) ...
  T q;
};

template <typename T>
inline __host__ __device__ EsirkConst<T> make_esirk_const(double q, double xmin, double ymin, double zmin, double dt,
                                       double dx, double dy, double dz) {
  EsirkConst<T> c;
  const double dxi = 1.0 / dx, dyi = 1.0 / dy, dzi = 1.0 / dz;  // esirkepov_3d.F90:3318-3329
  c.xmin = (T)xmin; c.ymin = (T)ymin; c.zmin = (T)zmin;
  c.dxi = (T)dxi; c.dyi = (T)dyi; c.dzi = (T)dzi;
  c.dtsdx0 = (T)(dt * dxi); c.dtsdy0 = (T)(dt * dyi); c.dtsdz0 = (T)(dt * dzi);
  c.invdtdx = (T)(1.0 / (dt * dy * dz));
  c.invdtdy = (T)(1.0 / (dt * dx * dz));
  c.invdtdz = (T)(1.0 / (dt * dx * dy));
  c.q = (T)q;
  return c;
}

template <typename T> struct EsirkepovIn {
  T x, y, z, xo, yo, zo, wqx, wqy, wqz;
};

// positions relative to the array origin, old positions and weights
template <int N, typename T>
PXR_DEV void esirkepov_prepare(EsirkepovIn<T> &in, T xp, T yp, T zp, T ux, T uy, T uz, T gi, T w,
                               const EsirkConst<T> &c) {
  in.x = (xp - c.xmin) * c.dxi;
  in.y = (yp - c.ymin) * c.dyi;
  in.z = (zp - c.zmin) * c.dzi;
  if (N == 1) {  // esirkepov_3d.F90:151-153
    in.xo = in.x - c.dtsdx0 * ux * gi;
    in.yo = in.y - c.dtsdy0 * uy * gi;
    in.zo = in.z - c.dtsdz0 * uz * gi;
  } else {  // esirkepov_3d.F90:3335-3343
    T vx = ux * gi, vy = uy * gi, vz = uz * gi;
    in.xo = in.x - c.dtsdx0 * vx;
    in.yo = in.y - c.dtsdy0 * vy;
    in.zo = in.z - c.dtsdz0 * vz;
  }
  T wq = c.q * w;
  in.wqx = wq * c.invdtdx;
  in.wqy = wq * c.invdtdy;
  in.wqz = wq * c.invdtdz;
}

// node range [lo, hi] touched by one particle; false if it moved more
// than one cell (the reference's shape arrays do not hold that case)
template <int N, typename T> PXR_DEV bool esirkepov_box(const EsirkepovIn<T> &in, int *lo, int *hi) {
  const T p[3] = {in.x, in.y, in.z}, o[3] = {in.xo, in.yo, in.zo};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int i0 = eindex<N, T>(p[a]), i1 = eindex<N, T>(o[a]);
    int d = i1 - i0;
    if (d < -1 || d > 1) return false;
    lo[a] = i0 + min(0, d) + ESup<N>::lo;
    hi[a] = i0 + max(0, d) + ESup<N>::hi;
  }
  return true;
}

// box-relative shapes along one axis: first node ibase, width N+1+|d|,
// new-position weights sn[], old-position weights so[] (zero outside)
template <int N, typename T> struct EAxis {
  int base, w;       // first node, number of nodes (N+1+|d|)
  T sn[N + 2], so[N + 2];
};

template <int N, typename T> PXR_DEV void esirkepov_axis(T x, T xo, EAxis<N, T> &ax) {
  const int i0 = eindex<N, T>(x), i1 = eindex<N, T>(xo);
  const int d = i1 - i0;  // -1, 0, 1 (checked by the caller)
  T wn[N + 1], wo[N + 1];
  eweights<N, T>(x, i0, wn);
  eweights<N, T>(xo, i1, wo);
  ax.base = i0 + min(0, d) + ESup<N>::lo;
  ax.w = N + 1 + (d < 0 ? -d : d);
#pragma unroll
  for (int r = 0; r < N + 2; r++) {
    // new shape starts at box index (d<0 ? 1 : 0); old at (d>0 ? 1 : 0)
    T a0 = (r <= N) ? wn[r < N + 1 ? r : N] : T(0);
    T a1 = (r >= 1) ? wn[r >= 1 ? r - 1 : 0] : T(0);
    ax.sn[r] = (d < 0) ? a1 : a0;
    T b0 = (r <= N) ? wo[r < N + 1 ? r : N] : T(0);
    T b1 = (r >= 1) ? wo[r >= 1 ? r - 1 : 0] : T(0);
    ax.so[r] = (d > 0) ? b1 : b0;
  }
}

// Accumulate one particle's current.  Acc::add(comp, i, j, k, value).
template <int N, typename T, typename Acc>
PXR_DEV void esirkepov_deposit(const EsirkepovIn<T> &in, const Acc &acc) {
  EAxis<N, T> X, Y, Z;
  esirkepov_axis<N, T>(in.x, in.xo, X);
  esirkepov_axis<N, T>(in.y, in.yo, Y);
  esirkepov_axis<N, T>(in.z, in.zo, Z);
  constexpr int W = N + 2;
  T dx_[W], dy_[W], dz_[W], cx[W], cy[W], cz[W];
#pragma unroll
  for (int r = 0; r < W; r++) {
    dx_[r] = X.so[r] - X.sn[r];
    dy_[r] = Y.so[r] - Y.sn[r];
    dz_[r] = Z.so[r] - Z.sn[r];
  }
  // prefix sums of dS along each axis, scaled by the charge factor
  T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
  for (int r = 0; r < W; r++) {
    sx += dx_[r];
    sy += dy_[r];
    sz += dz_[r];
    cx[r] = in.wqx * sx;
    cy[r] = in.wqy * sy;
    cz[r] = in.wqz * sz;
  }
  // transverse weights (esirkepov_3d.F90:3427-3450): for Jx
  // (S_y + dS_y/2) S_z + (S_y/2 + dS_y/3) dS_z, and likewise for Jy and Jz.
  // The two factors of each are formed once per node of their axis, and
  // each component runs its own loop nest, so every transverse weight is
  // evaluated once per pair of transverse nodes.
  const T third = Cst<T>::onethird;
  T hx[W], tx[W], hy[W], ty[W], hz[W], tz[W];
#pragma unroll
  for (int r = 0; r < W; r++) {
    hx[r] = X.sn[r] + T(0.5) * dx_[r];
    tx[r] = T(0.5) * X.sn[r] + third * dx_[r];
    hy[r] = Y.sn[r] + T(0.5) * dy_[r];
    ty[r] = T(0.5) * Y.sn[r] + third * dy_[r];
    hz[r] = Z.sn[r] + T(0.5) * dz_[r];
    tz[r] = T(0.5) * Z.sn[r] + third * dz_[r];
  }
  // Every component runs over rows of x nodes, so that a row's address is
  // formed once and its nodes are constant offsets from it.
#pragma unroll
  for (int rk = 0; rk < W; rk++) {
    if (rk >= Z.w) continue;
#pragma unroll
    for (int rj = 0; rj < W; rj++) {
      if (rj >= Y.w) continue;
      const T wx = hy[rj] * Z.sn[rk] + ty[rj] * dz_[rk];
      const auto row = acc.row(0, X.base, Y.base + rj, Z.base + rk);
#pragma unroll
      for (int ri = 0; ri < W - 1; ri++)
        if (ri < X.w - 1) acc.add_row(row, ri, cx[ri] * wx);
    }
  }
#pragma unroll
  for (int rk = 0; rk < W; rk++) {
    if (rk >= Z.w) continue;
    T wy[W];
#pragma unroll
    for (int ri = 0; ri < W; ri++) wy[ri] = hz[rk] * X.sn[ri] + tz[rk] * dx_[ri];
#pragma unroll
    for (int rj = 0; rj < W - 1; rj++) {
      if (rj >= Y.w - 1) continue;
      const auto row = acc.row(1, X.base, Y.base + rj, Z.base + rk);
#pragma unroll
      for (int ri = 0; ri < W; ri++)
        if (ri < X.w) acc.add_row(row, ri, cy[rj] * wy[ri]);
    }
  }
#pragma unroll
  for (int rj = 0; rj < W; rj++) {
    if (rj >= Y.w) continue;
    T wz[W];
#pragma unroll
    for (int ri = 0; ri < W; ri++) wz[ri] = hx[ri] * Y.sn[rj] + tx[ri] * dy_[rj];
#pragma unroll
    for (int rk = 0; rk < W - 1; rk++) {
      if (rk >= Z.w - 1) continue;
      const auto row = acc.row(2, X.base, Y.base + rj, Z.base + rk);
#pragma unroll
      for (int ri = 0; ri < W; ri++)
        if (ri < X.w) acc.add_row(row, ri, cz[rk] * wz[ri]);
    }
  }
}

// ---------------------------------------------------------------------------
// direct deposition
// ---------------------------------------------------------------------------
template <typename T> struct DirectIn {
  int j[3], j0[3];
  T s[3][4], s0[3][4];
  T wq[3];
};

template <int N, typename T>
PXR_DEV void direct_prepare(DirectIn<T> &in, T xp, T yp, T zp, T ux, T uy, T uz, T gi, T w, T q,
                            T xmin, T ymin, T zmin, T dt, T dx, T dy, T dz, T sft) {
  const T dxi = T(1) / dx, dyi = T(1) / dy, dzi = T(1) / dz;
  const T invvol = dxi * dyi * dzi;
  const T p[3] = {(xp - xmin) * dxi, (yp - ymin) * dyi, (zp - zmin) * dzi};
  const T v[3] = {ux * gi, uy * gi, uz * gi};
  const T h[3] = {T(0.5) * dt * dxi, T(0.5) * dt * dyi, T(0.5) * dt * dzi};
  const T wq = q * w * invvol;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    in.wq[a] = wq * v[a];
    const T xm = p[a] - h[a] * v[a];
    in.j[a] = dshape<N, T>(xm, T(0), in.s[a]);
    in.j0[a] = dshape<N, T>(xm, sft, in.s0[a]);
  }
}

template <int N, typename T> PXR_DEV void direct_box(const DirectIn<T> &in, int *lo, int *hi) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    lo[a] = min(in.j[a], in.j0[a]);
    hi[a] = max(in.j[a], in.j0[a]) + N;
  }
}

template <int N, typename T, typename Acc>
PXR_DEV void direct_deposit(const DirectIn<T> &in, const Acc &acc) {
#pragma unroll
  for (int c = 0; c <= N; c++)
#pragma unroll
    for (int b = 0; b <= N; b++)
#pragma unroll
      for (int a = 0; a <= N; a++) {
        acc.add(0, in.j0[0] + a, in.j[1] + b, in.j[2] + c, in.s0[0][a] * in.s[1][b] * in.s[2][c] * in.wq[0]);
        acc.add(1, in.j[0] + a, in.j0[1] + b, in.j[2] + c, in.s[0][a] * in.s0[1][b] * in.s[2][c] * in.wq[1]);
        acc.add(2, in.j[0] + a, in.j[1] + b, in.j0[2] + c, in.s[0][a] * in.s[1][b] * in.s0[2][c] * in.wq[2]);
      }
}

// ---------------------------------------------------------------------------
// wavefront ballot ranking
// ---------------------------------------------------------------------------
// For every active lane: rank among the active lanes of its wave holding the
// same key and a lower lane id, the number of such lanes (cnt) and the
// lowest of them (leader).  One ballot per distinct key in the wave.
PXR_DEV void wave_peers(bool active, int key, int &rank, int &cnt, int &leader) {
  const int lane = __lane_id();
  unsigned long long todo = __ballot(active);
  unsigned long long mine = 0;
  while (todo) {
    const int l = __ffsll((unsigned long long)todo) - 1;
    const int lk = __shfl(key, l);
    const bool same = active && key == lk;
    const unsigned long long m = __ballot(same);
    if (same) mine = m;
    todo &= ~m;
  }
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  rank = __popcll(mine & below);
  cnt = __popcll(mine);
  leader = mine ? __ffsll((unsigned long long)mine) - 1 : lane;
}

// Runs of equal keys on neighbouring active lanes (keys that are mostly
// sorted along the lanes): head is the lowest lane of this lane's run, cnt
// the run's length on its head lane and 0 elsewhere.  A run ends at a key
// change or an inactive lane.  Two ballots, whatever the number of keys.
PXR_DEV void wave_runs(int key, int &head, int &cnt) {
  const int lane = __lane_id();
  const unsigned long long act = __ballot(true);
  const int prev = __shfl(key, (lane + 63) & 63);
  const bool h = lane == 0 || !((act >> (lane - 1)) & 1ull) || prev != key;
  const unsigned long long heads = __ballot(h);
  const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane (all for lane 63)
  head = 63 - __clzll((long long)(heads & upto));
  // boundaries strictly above this lane
  const unsigned long long stop = (heads | ~act) & ~upto;
  const int next = stop ? __ffsll((unsigned long long)stop) - 1 : 64;
  cnt = h ? next - lane : 0;
}

// Block index for 1-D grids whose neighbouring blocks touch neighbouring
// memory: workgroups are dispatched to the 8 XCDs round-robin, so block b
// runs on XCD b % 8; this maps the blocks of each XCD to one contiguous
// range of logical blocks (a bijection on [0, nblocks)), so that lines two
// neighbouring blocks share stay in one XCD's L2.
PXR_DEV i64 xcd_block(i64 b, i64 nblocks) {
  const i64 q = nblocks / 8, r = nblocks % 8, x = b % 8, j = b / 8;
  return x * q + (x < r ? x : r) + j;
}

}  // namespace pxr
