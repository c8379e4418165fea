// pxr_mixed.h -- mixed-precision building blocks of the tiled PIC step.
//
// Positions stay FP64.  Cell indices and the offset of a particle inside its
// cell are computed in FP64 from them; the shape weights, the gather, the
// push and the per-particle current are FP32.  The Esirkepov current is a
// difference of shapes at the old and new positions: for a particle that
// stays on the same node index along an axis (almost all of them) the
// difference is evaluated in factored form, (a - b) Q(a, b), with the
// displacement a - b taken from FP64, so it keeps FP32 relative precision
// however small the displacement; an axis on which the particle changes
// node index is evaluated in FP64.
#pragma once
#include <type_traits>

#include "pxr_kernels.h"

namespace pxr {

// nodal and staggered shapes: index and offset from the FP64 position,
// weights in FP32 (same index rules as shape_nodal / shape_stag)
template <int N> PXR_DEV int shape_nodal_mx(double x, float *w) {
  if (N == 1) {
    const int j = f_floor(x);
    w_lin<float>((float)(x - (double)j), w);
    return j;
  }
  if (N == 2) {
    const int j = f_nint(x);
    w_quad<float>((float)(x - (double)j), w);
    return j - 1;
  }
  const int j = f_floor(x);
  w_cubic<float>((float)(x - (double)j), w);
  return j - 1;
}

template <int M> PXR_DEV int shape_stag_mx(double x, double s, float *w) {
  if (M == 0) {
    w[0] = 1.0f;
    return f_floor((x + 0.5) - s);
  }
  if (M == 1) {
    const int j0 = f_floor(x - s);
    w_lin<float>((float)((x - s) - (double)j0), w);
    return j0;
  }
  if (M == 2) {
    const int j0 = f_floor((x + 0.5) - s);
    w_quad<float>((float)((x - s) - (double)j0), w);
    return j0 - 1;
  }
  const int j0 = f_floor(x - s);
  w_cubic<float>((float)((x - s) - (double)j0), w);
  return j0 - 1;
}

template <int N, int M>
PXR_DEV void gather_stencil_mx(GatherStencil<N, M, float> &st, double x, double y, double z,
                               double s) {
  st.j = shape_nodal_mx<N>(x, st.sx);
  st.k = shape_nodal_mx<N>(y, st.sy);
  st.l = shape_nodal_mx<N>(z, st.sz);
  st.j0 = shape_stag_mx<M>(x, s, st.sx0);
  st.k0 = shape_stag_mx<M>(y, s, st.sy0);
  st.l0 = shape_stag_mx<M>(z, s, st.sz0);
}

// dw[r] = w_r(a) - w_r(b) for the order-N weights at offsets a and b of the
// same node index, with d = a - b: every difference is d times a
// polynomial in a and b (exact factorisation of the B-spline pieces)
template <int N> PXR_DEV void wdiff(float a, float b, float d, float *dw) {
  if (N == 1) {
    dw[0] = -d;
    dw[1] = d;
  } else if (N == 2) {
    dw[0] = -0.5f * d * (1.0f - a - b);
    dw[1] = -d * (a + b);
    dw[2] = 0.5f * d * (1.0f + a + b);
  } else {
    const float s2 = a * a + a * b + b * b;
    const float ga = 1.0f - a, gb = 1.0f - b;
    const float t2 = ga * ga + ga * gb + gb * gb;
    dw[0] = -d * t2 * (1.0f / 6.0f);
    dw[1] = d * (0.5f * s2 - (a + b));
    dw[2] = -d * (0.5f * t2 - (ga + gb));
    dw[3] = d * s2 * (1.0f / 6.0f);
  }
}

// Esirkepov shapes along one axis in the box of the particle: first node,
// width N+1+|d|, new-position weights sn[] and differences ds[] = old - new
template <int N> struct EAxisMx {
  int base, w;
  float sn[N + 2], ds[N + 2];
};

// u, uo: new and old position in cells (FP64)
template <int N> PXR_DEV void esirkepov_axis_mx(double u, double uo, EAxisMx<N> &ax) {
  const int i0 = eindex<N, double>(u), i1 = eindex<N, double>(uo);
  const int d = i1 - i0;  // -1, 0, 1 (checked by the caller)
  ax.base = i0 + min(0, d) + ESup<N>::lo;
  ax.w = N + 1 + (d < 0 ? -d : d);
  if (d == 0) {
    float wn[N + 1], dw[N + 1];
    const float b = (float)(u - (double)i0), a = (float)(uo - (double)i0);
    const float dd = (float)(uo - u);
    if (N == 1) w_lin<float>(b, wn);
    else if (N == 2) w_quad<float>(b, wn);
    else w_cubic<float>(b, wn);
    wdiff<N>(a, b, dd, dw);
#pragma unroll
    for (int r = 0; r < N + 2; r++) {
      ax.sn[r] = r <= N ? wn[r < N + 1 ? r : N] : 0.0f;
      ax.ds[r] = r <= N ? dw[r < N + 1 ? r : N] : 0.0f;
    }
  } else {
    // the node index changes along this axis: both shapes in FP64
    double wn[N + 1], wo[N + 1];
    eweights<N, double>(u, i0, wn);
    eweights<N, double>(uo, i1, wo);
#pragma unroll
    for (int r = 0; r < N + 2; r++) {
      const double a0 = (r <= N) ? wn[r < N + 1 ? r : N] : 0.0;
      const double a1 = (r >= 1) ? wn[r >= 1 ? r - 1 : 0] : 0.0;
      const double sn = (d < 0) ? a1 : a0;
      const double b0 = (r <= N) ? wo[r < N + 1 ? r : N] : 0.0;
      const double b1 = (r >= 1) ? wo[r >= 1 ? r - 1 : 0] : 0.0;
      const double so = (d > 0) ? b1 : b0;
      ax.sn[r] = (float)sn;
      ax.ds[r] = (float)(so - sn);
    }
  }
}

// one particle's Esirkepov current (same running sums as esirkepov_deposit)
template <int N, typename Acc>
PXR_DEV void esirkepov_accum_mx(const EAxisMx<N> &X, const EAxisMx<N> &Y, const EAxisMx<N> &Z,
                                float wqx, float wqy, float wqz, const Acc &acc) {
  constexpr int W = N + 2;
  float cx[W], cy[W], cz[W];
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
  for (int r = 0; r < W; r++) {
    sx += X.ds[r];
    sy += Y.ds[r];
    sz += Z.ds[r];
    cx[r] = wqx * sx;
    cy[r] = wqy * sy;
    cz[r] = wqz * sz;
  }
  const float third = 1.0f / 3.0f;
#pragma unroll
  for (int rk = 0; rk < W; rk++) {
    if (rk >= Z.w) continue;
#pragma unroll
    for (int rj = 0; rj < W; rj++) {
      if (rj >= Y.w) continue;
      const float wx = (Y.sn[rj] + 0.5f * Y.ds[rj]) * Z.sn[rk] +
                       (0.5f * Y.sn[rj] + third * Y.ds[rj]) * Z.ds[rk];
#pragma unroll
      for (int ri = 0; ri < W; ri++) {
        if (ri >= X.w) continue;
        const int i = X.base + ri, j = Y.base + rj, k = Z.base + rk;
        if (ri < X.w - 1) acc.add(0, i, j, k, cx[ri] * wx);
        if (rj < Y.w - 1) {
          const float wy = (Z.sn[rk] + 0.5f * Z.ds[rk]) * X.sn[ri] +
                           (0.5f * Z.sn[rk] + third * Z.ds[rk]) * X.ds[ri];
          acc.add(1, i, j, k, cy[rj] * wy);
        }
        if (rk < Z.w - 1) {
          const float wz = (X.sn[ri] + 0.5f * X.ds[ri]) * Y.sn[rj] +
                           (0.5f * X.sn[ri] + third * X.ds[ri]) * Y.ds[rj];
          acc.add(2, i, j, k, cz[rk] * wz);
        }
      }
    }
  }
}

// direct deposition: shapes at the mid position (FP64 index and offset)
template <int N> PXR_DEV int dshape_mx(double xm, double s, float *w) {
  if (N == 1) {
    const int j = f_floor(xm - s);
    w_lin<float>((float)((xm - s) - (double)j), w);
    return j;
  }
  if (N == 2) {
    const int j = f_nint(xm - s);
    w_quad<float>((float)((xm - s) - (double)j), w);
    return j - 1;
  }
  const int j = f_floor(xm - s);
  w_cubic<float>((float)((xm - s) - (double)j), w);
  return j - 1;
}

// pn: new position in cells (FP64); v: velocity; h = dt/(2 dx) per axis;
// wq = q w / volume
template <int N>
PXR_DEV void direct_prepare_mx(DirectIn<float> &in, const double *pn, const float *v,
                               const double *h, float wq, double sft) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    in.wq[a] = wq * v[a];
    const double xm = pn[a] - h[a] * (double)v[a];
    in.j[a] = dshape_mx<N>(xm, 0.0, in.s[a]);
    in.j0[a] = dshape_mx<N>(xm, sft, in.s0[a]);
  }
}

// Fixed-point value of v: round(v * scale) as a 64-bit two's-complement
// integer, for |v * scale| < 2^51.  Adding 1.5 * 2^52 rounds v * scale to an
// integer in the low mantissa bits (the ulp there is 1); subtracting the
// constant's bit pattern leaves that integer.
PXR_DEV unsigned long long fixed_point(double v, double scale) {
  const double magic = 6755399441055744.0;  // 1.5 * 2^52
  const double t = fma(v, scale, magic);
  return (unsigned long long)(__double_as_longlong(t) - __double_as_longlong(magic));
}

// LDS accumulators of the current block of a mixed-precision species,
// selected at compile time (JACC):
//   0  FP32 atomics;
//   1  FP32 atomics compensated: the rounding error of every add, recovered
//      from the returned old value by TwoSum, goes to a second FP32 array;
//   2  FP64 atomics on the FP32 contributions;
//   3  64-bit integer atomics on the contributions in fixed point (scale a
//      power of two per work item): exact, order-independent sums.
template <int JACC> struct LdsAccMx {
  typedef typename std::conditional<
      JACC == 3, unsigned long long,
      typename std::conditional<JACC == 2, double, float>::type>::type AT;
  AT *p[3];
  float *e[3];
  int i0, j0, k0, sy, sz;
  double scale;  // JACC 3
  PXR_DEV void add(int c, int i, int j, int k, float v) const {
    const int o = (i - i0) + sy * (j - j0) + sz * (k - k0);
    if constexpr (JACC == 3) {
      atomicAdd(p[c] + o, fixed_point((double)v, scale));
    } else if constexpr (JACC == 1) {
      const float old = atomicAdd(p[c] + o, v);
      const float s = old + v;
      const float bp = s - old;
      const float err = (old - (s - bp)) + (v - bp);
      atomicAdd(e[c] + o, err);
    } else {
      atomicAdd(p[c] + o, (AT)v);
    }
  }
};

// global FP64 current from FP32 contributions
struct GAccF {
  double *p[3];
  GridGeom g;
  PXR_DEV void add(int c, int i, int j, int k, float v) const {
    unsafeAtomicAdd(p[c] + g.idx(i, j, k), (double)v);
  }
};

// FP64 field arrays read as FP32
struct FieldViewF {
  const double *p;  // points at node (0,0,0)
  i64 sy, sz;
  PXR_DEV float at(int i, int j, int k) const { return (float)p[i + sy * j + sz * k]; }
};

}  // namespace pxr
