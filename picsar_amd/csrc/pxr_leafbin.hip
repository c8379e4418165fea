// pxr_leafbin.hip -- the leaf current and charge deposition (depose_jxjyjz
// with the Esirkepov current, pxr_depose_rho_n) for large particle sets,
// summed block by block in LDS.
//
// A caller of the leaf routines hands over particles in any order, anywhere
// in the grid arrays.  Adding every contribution to global memory costs one
// FP64 L2 atomic per node and particle (144 for an order-3 Esirkepov
// particle that keeps its nodes), and those atomics bound the kernel.  Here
// the particles are first counted into cubic blocks of B^3 nodes (by the
// node floor(x) of their new position), ordered by block with a counting
// sort of their indices, and cut into work items of at most `chunk`
// particles of one block.  A workgroup sums its work item's contributions
// into an LDS copy of the block with its halo (the nodes that the shapes
// and one cell of motion reach) and adds the block to the grid array once
// per non-zero node.
//
// The contributions are the same arithmetic as the global-atomic kernels
// (esirkepov_deposit, the nodal shapes of pxr_depose_rho_n); only the order
// of the FP64 sums differs.  Nothing synchronises with the host, so the
// sequence can be captured in a HIP graph once its scratch is allocated.
#include <algorithm>

#include "pxr_leafbin.h"
#include "pxr_runtime.h"

namespace pxr {

// Blocks of B^3 nodes over the padded array; block (bx, by, bz) holds the
// particles whose node floor(x) + guard lies in [bx B, bx B + B).
struct BinGeom {
  int nb[3];   // blocks per axis
  int gd[3];   // guard widths
  int na[3];   // padded node counts
  i64 nblocks;
};

static BinGeom make_bin_geom(const GridGeom &g, int B) {
  BinGeom b;
  b.gd[0] = g.gx; b.gd[1] = g.gy; b.gd[2] = g.gz;
  b.na[0] = g.nx + 1 + 2 * g.gx;
  b.na[1] = g.ny + 1 + 2 * g.gy;
  b.na[2] = g.nz + 1 + 2 * g.gz;
  for (int d = 0; d < 3; d++) b.nb[d] = (b.na[d] + B - 1) / B;
  b.nblocks = (i64)b.nb[0] * b.nb[1] * b.nb[2];
  return b;
}

// block index along one axis of normalised position x: floor(x) clamped to
// the array (a particle outside it is rejected by the deposit kernel)
template <int B> PXR_DEV int block_axis(double x, int g, int na) {
  double f = floor(x);
  f = fmin(fmax(f, -(double)g), (double)(na - 1 - g));
  return ((int)f + g) / B;
}

__global__ __launch_bounds__(256) void k_lb_fill(int *p, i64 n) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Runs of equal keys among a workgroup's particles (keys in skey[0..n),
// every lane of the workgroup calls): the first lane of lane t's run.  A
// wave finds the nearest run start at or before each lane from the ballot
// of run starts; lanes with none in their wave take the last one of the
// waves before.
PXR_DEV int lb_run_start(const int *skey, int *wlast, int t, int n) {
  const int lane = t & 63, w = t >> 6;
  const bool head = t < n && (t == 0 || skey[t - 1] != skey[t]);
  const unsigned long long H = __ballot(head);
  if (lane == 0) wlast[w] = H ? (w << 6) + 63 - __clzll((long long)H) : -1;
  __syncthreads();
  const unsigned long long le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  const unsigned long long hb = H & le;
  if (hb) return (w << 6) + 63 - __clzll((long long)hb);
  int st = -1;
  for (int v = 0; v < w; v++) st = max(st, wlast[v]);
  return st;
}

// the last lane of a run of equal keys
PXR_DEV bool lb_run_end(const int *skey, int t, int n) {
  return t < n && (t == n - 1 || skey[t + 1] != skey[t]);
}

// Per particle its block, and counts per block.  The counts are kept in
// `copies` counters per block, workgroup w adding to copy w % copies, and
// each run of equal keys in a workgroup adds its length with one atomic:
// cell-ordered input, where consecutive workgroups meet the same blocks,
// then spreads its few atomics over many addresses.
template <int B>
__global__ __launch_bounds__(1024) void k_lb_count(i64 np, const double *xp, const double *yp,
                                                   const double *zp, double xmin, double ymin,
                                                   double zmin, double dxi, double dyi,
                                                   double dzi, int l4symtry, BinGeom bg,
                                                   int copies, int *key, int *cnt) {
  __shared__ int skey[1024], wlast[16];
  const int t = threadIdx.x;
  const i64 base = (i64)blockIdx.x * 1024, ip = base + t;
  const int n = (int)min((i64)1024, np - base);
  int b = -1;
  if (ip < np) {
    double x = (xp[ip] - xmin) * dxi, y = (yp[ip] - ymin) * dyi;
    const double z = (zp[ip] - zmin) * dzi;
    if (l4symtry) {
      x = fabs(x);
      y = fabs(y);
    }
    b = block_axis<B>(x, bg.gd[0], bg.na[0]) +
        bg.nb[0] * (block_axis<B>(y, bg.gd[1], bg.na[1]) +
                    bg.nb[1] * block_axis<B>(z, bg.gd[2], bg.na[2]));
    key[ip] = b;
  }
  skey[t] = b;
  __syncthreads();
  const int st = lb_run_start(skey, wlast, t, n);
  if (lb_run_end(skey, t, n)) atomicAdd(cnt + (i64)b * copies + blockIdx.x % copies, t - st + 1);
}

// One workgroup: exclusive scan of the counts, block by block and copy by
// copy, into the copies' first slots (cursor), and the work items (block,
// first slot, count) of at most `chunk` particles of one block.  Each
// thread owns a contiguous range of blocks.
__global__ __launch_bounds__(1024) void k_lb_plan(const int *cnt, i64 nblocks, int copies,
                                                  int chunk, int *cursor, int4 *items,
                                                  int *nitems) {
  __shared__ long long sp[1024], si[1024];
  const int t = threadIdx.x;
  const i64 per = (nblocks + 1023) / 1024;
  const i64 b0 = min(nblocks, (i64)t * per), b1 = min(nblocks, b0 + per);
  long long np = 0, ni = 0;
  for (i64 b = b0; b < b1; b++) {
    int c = 0;
    for (int q = 0; q < copies; q++) c += cnt[b * copies + q];
    np += c;
    ni += (c + chunk - 1) / chunk;
  }
  sp[t] = np;
  si[t] = ni;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const long long a = t >= off ? sp[t - off] : 0, c = t >= off ? si[t - off] : 0;
    __syncthreads();
    sp[t] += a;
    si[t] += c;
    __syncthreads();
  }
  long long pos = sp[t] - np, it = si[t] - ni;
  for (i64 b = b0; b < b1; b++) {
    const long long first = pos;
    for (int q = 0; q < copies; q++) {
      cursor[b * copies + q] = (int)pos;
      pos += cnt[b * copies + q];
    }
    const int c = (int)(pos - first);
    for (int k = 0; k < c; k += chunk)
      items[it++] = make_int4((int)b, (int)(first + k), min(chunk, c - k), 0);
  }
  if (t == 1023) *nitems = (int)si[1023];
}

// Particle indices in block order: each run of equal keys in a workgroup
// claims its slots with one atomic on its copy's cursor (the workgroups
// and copies of k_lb_count), so a block's slots keep the input order
// within each run.
__global__ __launch_bounds__(1024) void k_lb_scatter(i64 np, const int *key, int copies,
                                                     int *cursor, int *perm) {
  __shared__ int skey[1024], sbase[1024], wlast[16];
  const int t = threadIdx.x;
  const i64 base = (i64)blockIdx.x * 1024, ip = base + t;
  const int n = (int)min((i64)1024, np - base);
  const int b = ip < np ? key[ip] : -1;
  skey[t] = b;
  __syncthreads();
  const int st = lb_run_start(skey, wlast, t, n);
  if (lb_run_end(skey, t, n))
    sbase[st] = atomicAdd(cursor + (i64)b * copies + blockIdx.x % copies, t - st + 1);
  __syncthreads();
  if (ip < np) perm[sbase[st] + (t - st)] = (int)ip;
}

// LDS block of one work item: nodes [o0, o0 + L) per axis (grid indices)
struct LdsBlock {
  double *p;
  int o[3], L, nodes;
  PXR_DEV bool holds(const int *lo, const int *hi) const {
    return lo[0] >= o[0] && lo[1] >= o[1] && lo[2] >= o[2] && hi[0] < o[0] + L &&
           hi[1] < o[1] + L && hi[2] < o[2] + L;
  }
};

// LDS accumulator of the current (FP64 LDS atomics), component c at p + c*nodes
struct LdsAccB {
  double *p;
  int i0, j0, k0, sy, sz, nodes;
  PXR_DEV double *row(int c, int i, int j, int k) const {
    return p + c * nodes + ((i - i0) + sy * (j - j0) + sz * (k - k0));
  }
  PXR_DEV void add_row(double *r, int o, double v) const { atomicAdd(r + o, v); }
  PXR_DEV void add(int c, int i, int j, int k, double v) const { atomicAdd(row(c, i, j, k), v); }
};

// the LDS block of block b: first node and edge; zeroed (after the
// previous work item's flush has read it)
template <int B, int HLO, int HHI>
PXR_DEV void lb_block(const BinGeom &bg, int b, LdsBlock &blk, double *lds, int ncomp) {
  const int bx = b % bg.nb[0], by = (b / bg.nb[0]) % bg.nb[1], bz = b / (bg.nb[0] * bg.nb[1]);
  blk.p = lds;
  blk.L = B + HHI - HLO;
  blk.nodes = blk.L * blk.L * blk.L;
  blk.o[0] = bx * B - bg.gd[0] + HLO;
  blk.o[1] = by * B - bg.gd[1] + HLO;
  blk.o[2] = bz * B - bg.gd[2] + HLO;
  __syncthreads();
  for (int i = threadIdx.x; i < ncomp * blk.nodes; i += blockDim.x) lds[i] = 0.0;
  __syncthreads();
}

// add the LDS block's non-zero nodes to the grid arrays
PXR_DEV void lb_flush(const LdsBlock &blk, const GridGeom &g, double *const *dst, int ncomp) {
  __syncthreads();
  const int L = blk.L;
  for (int t = threadIdx.x; t < blk.nodes; t += blockDim.x) {
    const int i = blk.o[0] + t % L, j = blk.o[1] + (t / L) % L, k = blk.o[2] + t / (L * L);
    if (!g.inside(i, j, k)) continue;  // nothing deposits outside the array
    const i64 ix = g.idx(i, j, k);
    for (int c = 0; c < ncomp; c++) {
      const double v = blk.p[c * blk.nodes + t];
      if (v != 0.0) unsafeAtomicAdd(dst[c] + ix, v);
    }
  }
}

// Lanes take particles S = ceil(n / 256) apart, so that in cell-ordered
// input the lanes of a wave sit in different cells.
PXR_DEV int lb_stride(int n) { return (n + (int)blockDim.x - 1) / (int)blockDim.x; }

// Esirkepov current: the box of an order-N particle lies within
// [floor(x) - 2, floor(x) + 3] (order 1: [-1, +2]) on each axis.
template <int N, int B>
__global__ __launch_bounds__(256) void k_lb_esirkepov(DepoArgs a, BinGeom bg, const int *perm,
                                                      const int4 *items, const int *nitems) {
  constexpr int HLO = (N == 1) ? -1 : -2, HHI = (N == 1) ? 2 : 3;
  extern __shared__ double lds[];
  const i64 o = a.g.idx(0, 0, 0);
  const GlobalAcc gacc{{a.jx + o, a.jy + o, a.jz + o}, a.g.NX, a.g.NX * a.g.NY};
  double *dst[3] = {a.jx, a.jy, a.jz};
  const int nit = *nitems;
  for (int w = blockIdx.x; w < nit; w += gridDim.x) {
  const int4 it = items[w];
  LdsBlock blk;
  lb_block<B, HLO, HHI>(bg, it.x, blk, lds, 3);
  const LdsAccB acc{lds, blk.o[0], blk.o[1], blk.o[2], blk.L, blk.L * blk.L, blk.nodes};
  const int n = it.z, S = lb_stride(n);
  for (int k = 0; k < S; k++) {
    const int p = (int)threadIdx.x * S + k;
    if (p >= n) break;
    const i64 ip = perm[it.y + p];
    EsirkepovIn<double> in;
    esirkepov_prepare<N, double>(in, a.xp[ip], a.yp[ip], a.zp[ip], a.ux[ip], a.uy[ip],
                                 a.uz[ip], a.gi[ip], a.w[ip], a.ec);
    int lo[3], hi[3];
    if (!esirkepov_box<N, double>(in, lo, hi) || !a.g.inside(lo[0], lo[1], lo[2]) ||
        !a.g.inside(hi[0], hi[1], hi[2])) {
      count_bad(a.bad);
      continue;
    }
    if (blk.holds(lo, hi)) esirkepov_deposit<N, double>(in, acc);
    else esirkepov_deposit<N, double>(in, gacc);  // not reached: the halo holds every box
  }
  lb_flush(blk, a.g, dst, 3);
  }
}

// charge density: nodes [floor(x) - 1, floor(x) + 2] (order 1: [0, 1])
template <int N, int B>
__global__ __launch_bounds__(256) void k_lb_rho(RhoArgs a, BinGeom bg, const int *perm,
                                                const int4 *items, const int *nitems) {
  constexpr int HLO = (N == 1) ? 0 : -1, HHI = (N == 1) ? 1 : 2;
  extern __shared__ double lds[];
  double *dst[1] = {a.rho};
  const int nit = *nitems;
  for (int w = blockIdx.x; w < nit; w += gridDim.x) {
  const int4 it = items[w];
  LdsBlock blk;
  lb_block<B, HLO, HHI>(bg, it.x, blk, lds, 1);
  const int L = blk.L;
  const int n = it.z, S = lb_stride(n);
  for (int k = 0; k < S; k++) {
    const int p = (int)threadIdx.x * S + k;
    if (p >= n) break;
    const i64 ip = perm[it.y + p];
    double x = (a.xp[ip] - a.xmin) * a.dxi, y = (a.yp[ip] - a.ymin) * a.dyi;
    const double z = (a.zp[ip] - a.zmin) * a.dzi;
    if (a.l4symtry) {
      x = fabs(x);
      y = fabs(y);
    }
    double sx[N + 1], sy[N + 1], sz[N + 1];
    const int j = shape_nodal<N, double>(x, sx);
    const int kk = shape_nodal<N, double>(y, sy);
    const int l = shape_nodal<N, double>(z, sz);
    if (!a.g.inside(j, kk, l) || !a.g.inside(j + N, kk + N, l + N)) {
      count_bad(a.bad);
      continue;
    }
    const double wq =
        a.per_particle_weight ? a.q * a.w[ip] * a.invvol : a.q * a.invvol * a.w[0];
    const int lo[3] = {j, kk, l}, hi[3] = {j + N, kk + N, l + N};
    const bool in_lds = blk.holds(lo, hi);
    double *r = in_lds ? lds + ((j - blk.o[0]) + L * ((kk - blk.o[1]) + L * (l - blk.o[2])))
                       : a.rho + a.g.idx(j, kk, l);
    const i64 sy_ = in_lds ? L : a.g.NX, sz_ = in_lds ? (i64)L * L : a.g.NX * a.g.NY;
#pragma unroll
    for (int c = 0; c <= N; c++)
#pragma unroll
      for (int b = 0; b <= N; b++)
#pragma unroll
        for (int e = 0; e <= N; e++) {
          const double v = sx[e] * sy[b] * sz[c] * wq;
          if (in_lds) atomicAdd(r + e + sy_ * b + sz_ * c, v);
          else unsafeAtomicAdd(r + e + sy_ * b + sz_ * c, v);
        }
  }
  lb_flush(blk, a.g, dst, 1);
  }
}

// counting sort of the particle indices by block; work items into *items
template <int B>
static int lb_sort(i64 np, const double *xp, const double *yp, const double *zp, double xmin,
                   double ymin, double zmin, double dxi, double dyi, double dzi, int l4symtry,
                   const GridGeom &g, int chunk, hipStream_t st, BinGeom &bg, int *&perm,
                   int4 *&items, int *&nitems, i64 &maxitems) {
  bg = make_bin_geom(g, B);
  maxitems = bg.nblocks + np / chunk + 1;
  if (np >= (i64)1 << 31 || bg.nblocks >= (i64)1 << 31 || maxitems >= (i64)1 << 31)
    return set_error(PXRHIP_ERR_ARG, "binned deposition: more than 2^31 particles or blocks");
  // counters per block: up to 16 copies while they stay below 2^18
  const int copies = (int)std::max<i64>(1, std::min<i64>(16, ((i64)1 << 18) / bg.nblocks));
  const i64 ncnt = bg.nblocks * copies;
  int *key = (int *)scratch(24, sizeof(int) * (size_t)np);
  perm = (int *)scratch(25, sizeof(int) * (size_t)np);
  int *cnt = (int *)scratch(26, sizeof(int) * (size_t)ncnt);
  int *cursor = (int *)scratch(27, sizeof(int) * (size_t)ncnt);
  items = (int4 *)scratch(28, sizeof(int4) * (size_t)maxitems);
  nitems = (int *)scratch(29, sizeof(int));
  if (!key || !perm || !cnt || !cursor || !items || !nitems)
    return set_error(PXRHIP_ERR_HIP, "binned deposition: out of device memory");
  hipLaunchKernelGGL(k_lb_fill, dim3(grid1d(ncnt, 256)), dim3(256), 0, st, cnt, ncnt);
  hipLaunchKernelGGL(k_lb_count<B>, dim3(grid1d(np, 1024)), dim3(1024), 0, st, np, xp, yp, zp,
                     xmin, ymin, zmin, dxi, dyi, dzi, l4symtry, bg, copies, key, cnt);
  hipLaunchKernelGGL(k_lb_plan, dim3(1), dim3(1024), 0, st, cnt, bg.nblocks, copies, chunk,
                     cursor, items, nitems);
  hipLaunchKernelGGL(k_lb_scatter, dim3(grid1d(np, 1024)), dim3(1024), 0, st, np, key, copies,
                     cursor, perm);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

constexpr int LB_EDGE = 8;  // nodes per block edge

// persistent workgroups: as many as the CUs hold at this LDS size, or one
// per work item if there are fewer
static unsigned lb_grid(i64 maxitems, size_t lds) {
  const i64 per_cu = (i64)(160 * 1024 / (lds + 1024));
  const i64 g = (i64)rt().num_cus * (per_cu < 1 ? 1 : per_cu);
  return (unsigned)(maxitems < g ? maxitems : g);
}

#define LB_LAUNCH(KERNEL, ARGS)                                                               \
  do {                                                                                       \
    PXR_HIP(hipFuncSetAttribute((const void *)&KERNEL,                                      \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));     \
    hipLaunchKernelGGL(KERNEL, grd, blk, lds, st, ARGS, bg, perm, items, nitems);             \
  } while (0)

static bool pays(i64 np, const GridGeom &g, int edge) {
  const Runtime &r = rt();
  return r.leaf_bin_min > 0 && np >= r.leaf_bin_min && !r.capturing &&
         np >= make_bin_geom(g, edge).nblocks;
}

bool leaf_binned_esirkepov_pays(i64 np, const GridGeom &g) { return pays(np, g, LB_EDGE); }
bool leaf_binned_rho_pays(i64 np, const GridGeom &g) { return pays(np, g, LB_EDGE); }

int leaf_binned_esirkepov(const DepoArgs &a, int order, hipStream_t st) {
  BinGeom bg;
  int *perm, *nitems;
  int4 *items;
  i64 maxitems;
  const EsirkConst<double> &c = a.ec;
  int r = lb_sort<LB_EDGE>(a.np, a.xp, a.yp, a.zp, c.xmin, c.ymin, c.zmin, c.dxi, c.dyi, c.dzi, 0,
                           a.g, rt().leaf_bin_chunk, st, bg, perm, items, nitems, maxitems);
  if (r) return r;
  const int L = LB_EDGE + (order == 1 ? 3 : 5);
  const size_t lds = sizeof(double) * 3 * (size_t)(L * L * L);
  const dim3 grd(lb_grid(maxitems, lds)), blk(256);
  if (order == 1) LB_LAUNCH((k_lb_esirkepov<1, LB_EDGE>), a);
  else if (order == 2) LB_LAUNCH((k_lb_esirkepov<2, LB_EDGE>), a);
  else LB_LAUNCH((k_lb_esirkepov<3, LB_EDGE>), a);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int leaf_binned_rho(const RhoArgs &a, int order, hipStream_t st) {
  constexpr int E = LB_EDGE;
  BinGeom bg;
  int *perm, *nitems;
  int4 *items;
  i64 maxitems;
  int r = lb_sort<E>(a.np, a.xp, a.yp, a.zp, a.xmin, a.ymin, a.zmin, a.dxi, a.dyi, a.dzi,
                     a.l4symtry, a.g, rt().leaf_bin_chunk, st, bg, perm, items, nitems, maxitems);
  if (r) return r;
  const int L = E + (order == 1 ? 1 : 3);
  const size_t lds = sizeof(double) * (size_t)(L * L * L);
  const dim3 grd(lb_grid(maxitems, lds)), blk(256);
  if (order == 1) LB_LAUNCH((k_lb_rho<1, E>), a);
  else if (order == 2) LB_LAUNCH((k_lb_rho<2, E>), a);
  else LB_LAUNCH((k_lb_rho<3, E>), a);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

}  // namespace pxr
