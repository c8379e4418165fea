// pxr_device.h -- device-side building blocks of the PIC hot path for gfx950.
//
// Shape factors, index rules and the per-particle gather / push / deposit
// arithmetic used by every kernel in this library.  FP64 code paths keep the
// reference's operation order and are compiled without FP contraction so that
// integer cell indices (floor/nint of normalised positions) are the ones the
// reference computes; summations that do not feed an index may use fma.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PXR_DEV __device__ __forceinline__
#define PXR_HD __host__ __device__ __forceinline__

namespace pxr {

typedef int64_t i64;

// Fortran NINT (round half away from zero) and FLOOR.
PXR_HD int f_nint(double x) { return (int)round(x); }
PXR_HD int f_floor(double x) { return (int)floor(x); }
PXR_DEV int f_nint(float x) { return (int)roundf(x); }
PXR_DEV int f_floor(float x) { return (int)floorf(x); }

template <typename T> struct Cst {
  static constexpr T onesixth = T(1.0 / 6.0);
  static constexpr T twothird = T(2.0 / 3.0);
  static constexpr T onethird = T(1.0 / 3.0);
};

// B-spline weights in the reference's evaluation order
// (field_gathering_o1_3d.F90:145-150, o2_3d.F90:147-158, o3_3d.F90:175-201).
template <typename T> PXR_DEV void w_lin(T f, T* w) {
  w[0] = T(1) - f;
  w[1] = f;
}
template <typename T> PXR_DEV void w_quad(T f, T* w) {
  T a = T(0.5) - f, b = T(0.5) + f;
  w[0] = T(0.5) * (a * a);
  w[1] = T(0.75) - f * f;
  w[2] = T(0.5) * (b * b);
}
template <typename T> PXR_DEV void w_cubic(T f, T* w) {
  T o = T(1) - f, f2 = f * f, o2 = o * o;
  w[0] = Cst<T>::onesixth * o2 * o;
  w[1] = Cst<T>::twothird - f2 * (T(1) - f * T(0.5));
  w[2] = Cst<T>::twothird - o2 * (T(1) - o * T(0.5));
  w[3] = Cst<T>::onesixth * f2 * f;
}

// Node-centred shape of order N (gather and charge): floor for N = 1, 3,
// nint for N = 2.  Returns the first node, w[0..N].
template <int N, typename T> PXR_DEV int shape_nodal(T x, T* w) {
  if (N == 1) { int j = f_floor(x); w_lin<T>(x - (T)j, w); return j; }
  if (N == 2) { int j = f_nint(x); w_quad<T>(x - (T)j, w); return j - 1; }
  int j = f_floor(x); w_cubic<T>(x - (T)j, w); return j - 1;
}

// Staggered shape of order M (0..3) for the gather, shift s:
// M even -> floor((x+0.5)-s), M odd -> floor(x-s); weights of (x-s)-j0.
template <int M, typename T> PXR_DEV int shape_stag(T x, T s, T* w) {
  if (M == 0) { int j0 = f_floor((x + T(0.5)) - s); w[0] = T(1); return j0; }
  if (M == 1) { int j0 = f_floor(x - s); w_lin<T>((x - s) - (T)j0, w); return j0; }
  if (M == 2) { int j0 = f_floor((x + T(0.5)) - s); w_quad<T>((x - s) - (T)j0, w); return j0 - 1; }
  int j0 = f_floor(x - s); w_cubic<T>((x - s) - (T)j0, w); return j0 - 1;
}

// Direct deposition shapes at xmid (direct_current_deposition_3d.F90):
// full order N in both node sets; nint for N = 2 on both.
template <int N, typename T> PXR_DEV int dshape(T xm, T s, T* w) {
  if (N == 1) { int j = f_floor(xm - s); w_lin<T>((xm - s) - (T)j, w); return j; }
  if (N == 2) { int j = f_nint(xm - s); w_quad<T>((xm - s) - (T)j, w); return j - 1; }
  int j = f_floor(xm - s); w_cubic<T>((xm - s) - (T)j, w); return j - 1;
}

// Esirkepov node index of a position: floor for N = 1, 3, nint for N = 2.
template <int N, typename T> PXR_DEV int eindex(T x) {
  return N == 2 ? f_nint(x) : f_floor(x);
}
// Esirkepov weights of order N at position x relative to its own node
// index j = eindex(x); first weight at node j + elo<N>.
template <int N> struct ESup {
  static constexpr int lo = (N == 1) ? 0 : -1;   // first node of the support
  static constexpr int hi = (N == 3) ? 2 : 1;    // last node of the support
};
template <int N, typename T> PXR_DEV void eweights(T x, int j, T* w) {
  if (N == 1) w_lin<T>(x - (T)j, w);
  else if (N == 2) w_quad<T>(x - (T)j, w);
  else w_cubic<T>(x - (T)j, w);
}

// Geometry of one Fortran-ordered array with bounds
// (-gx:nx+gx, -gy:ny+gy, -gz:nz+gz).
struct GridGeom {
  int nx, ny, nz, gx, gy, gz;
  i64 NX, NY;  // padded extents in x and y
  PXR_DEV i64 idx(int i, int j, int k) const {
    return (i64)(i + gx) + NX * ((i64)(j + gy) + NY * (i64)(k + gz));
  }
  PXR_DEV bool inside(int i, int j, int k) const {
    return i >= -gx && i <= nx + gx && j >= -gy && j <= ny + gy && k >= -gz && k <= nz + gz;
  }
};

// Device-wide error counters (out-of-bounds particles and similar).
struct ErrCounters {
  unsigned long long* bad;  // [0]: particles skipped, [1]: other
};

PXR_DEV void count_bad(unsigned long long* c) {
  if (c) atomicAdd(c, 1ull);
}

}  // namespace pxr
