// pxr_tiles.hip -- the tiled step on device-resident particles.
//
// Work of the reference's managers, done per species over all tiles:
//   * pxrhip_tiles_gather_push_deposit: gather + push + position push
//     (field_gathering_plus_particle_pusher_cacheblock_sub,
//     particle_pusher_manager_3d.F90:431-672) fused with the current
//     deposition of the pushed particles (pxrdepose_currents_on_grid_jxjyjz,
//     current_deposition_manager_3d.F90:224-1830).  A persistent kernel pulls
//     chunks of cell-sorted particles from a work queue; for each chunk it
//     stages the six field blocks its stencils touch in LDS, gathers from
//     LDS, pushes in registers, accumulates the chunk's current block in LDS
//     and adds that block to the global current arrays once.
//   * pxrhip_tiles_rebin_sort: periodic wrap + tile re-binning
//     (particle_boundaries.F90:350-426, :1040-1300) and the per-tile cell
//     sort (sorting.F90:106-361) as one counting sort over (tile, cell) bins.
#include <limits.h>
#include <string.h>

#include <vector>

#include "pxr_device.h"
#include "pxr_fused.h"
#include "pxr_kernels.h"
#include "pxr_runtime.h"

using namespace pxr;

namespace pxr {

// ---------------------------------------------------------------------------
// tile geometry (host), following set_tile_split_for_species
// (tiling.F90:94-271) axis by axis
// ---------------------------------------------------------------------------
struct AxisSplit {
  int nt = 1, n0 = 1;                   // tiles, nodes per tile (first tile)
  std::vector<double> tmin, tmax;       // particle membership [tmin, tmax)
};

static int split_axis(int64_t ncells, int64_t nt_in, double xmin, double dx, AxisSplit &a) {
  const int64_t ngrid = ncells + 1;  // nodes (mpi_routines.F90:1505-1510)
  int64_t nt = nt_in;
  int64_t n0 = ngrid / nt;
  if (n0 < 4) return set_error(PXRHIP_ERR_ARG, "fewer than 4 nodes per tile (the reference falls back to 1 tile)");
  const int64_t n0_last = n0 + (ngrid - n0 * nt);
  const double xmax = xmin + (double)ncells * dx;
  a.nt = (int)nt;
  a.tmin.assign(nt, 0.0);
  a.tmax.assign(nt, 0.0);
  for (int64_t i = 0; i < nt; i++) {
    double gmin = 0, gmax = 0, tmin = 0, tmax = 0;
    int64_t ngt = 0;
    if (i == 0) {
      ngt = n0;
      gmin = xmin;
      gmax = gmin + (double)(n0 - 1) * dx;
      tmin = gmin;
      tmax = gmax + dx / 2.0;
    }
    if (i > 0 && i < nt - 1) {
      ngt = n0;
      gmin = xmin + (double)(i * n0) * dx;
      gmax = gmin + (double)(ngt - 1) * dx;
      tmin = gmin - dx / 2.0;
      tmax = gmax + dx / 2.0;
    }
    if (i == nt - 1) {
      ngt = n0_last;
      gmin = xmin + (double)(i * n0) * dx;
      gmax = gmin + (double)(ngt - 1) * dx;
      tmin = gmin - dx / 2.0;
      tmax = xmax;
    }
    a.tmin[i] = tmin;
    a.tmax[i] = tmax;
  }
  // particle_bcs_tiles reads the node count of tile 1 (particle_boundaries.F90:364)
  a.n0 = (int)(nt == 1 ? n0_last : n0);
  return PXRHIP_OK;
}

struct TileTable {
  pxrhip_grid key;
  bool valid = false;
  int ntiles = 0;
  int nt[3] = {1, 1, 1}, n0[3] = {1, 1, 1};
  int64_t nbins = 0, nrows = 0;
  int max_bins_tile = 0;
  int nxmax = 1, nzmax = 1;   // largest sorting-grid extents of a tile in x and z
  // device arrays, per tile
  double *sorg = nullptr;     // 3*ntiles: sorting-grid origin
  double *tbox = nullptr;     // 6*ntiles: particle membership [tmin, tmax) per axis
  int *n3 = nullptr;          // 3*ntiles: sorting cells per axis
  int64_t *bin_base = nullptr;  // ntiles+1
};

static TileTable g_tt;

static void free_table(TileTable &t) {
  if (t.sorg || t.n3 || t.bin_base || t.tbox) {
    (void)hipStreamSynchronize(rt().stream);
    rt().mem_gen++;  // recorded graphs that read the table are stale
  }
  if (t.sorg) (void)hipFree(t.sorg);
  if (t.tbox) (void)hipFree(t.tbox);
  if (t.n3) (void)hipFree(t.n3);
  if (t.bin_base) (void)hipFree(t.bin_base);
  t.sorg = nullptr;
  t.tbox = nullptr;
  t.n3 = nullptr;
  t.bin_base = nullptr;
  t.valid = false;
}

static int check_grid_desc(const pxrhip_grid *g) {
  if (!g) return set_error(PXRHIP_ERR_ARG, "null grid");
  if (g->nx < 1 || g->ny < 1 || g->nz < 1) return set_error(PXRHIP_ERR_ARG, "empty grid");
  if (g->ngx < 1 || g->ngy < 1 || g->ngz < 1) return set_error(PXRHIP_ERR_ARG, "guards must be >= 1");
  if (g->ntx < 1 || g->nty < 1 || g->ntz < 1) return set_error(PXRHIP_ERR_ARG, "tile counts must be >= 1");
  if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0)) return set_error(PXRHIP_ERR_ARG, "dx, dy, dz must be > 0");
  if (!(g->sorting_dx > 0) || !(g->sorting_dy > 0) || !(g->sorting_dz > 0))
    return set_error(PXRHIP_ERR_ARG, "sorting cell sizes must be > 0");
  for (int k = 0; k < 6; k++)
    if (g->pbound[k] < PXRHIP_PBOUND_PERIODIC || g->pbound[k] > PXRHIP_PBOUND_REINJECTING)
      return set_error(PXRHIP_ERR_ARG, "pbound[%d] = %lld is not a particle boundary (0..3)", k,
                       (long long)g->pbound[k]);
  if (g->partcom < 0 || g->partcom > 2)
    return set_error(PXRHIP_ERR_ARG, "partcom = %lld is not 0, 1 or 2", (long long)g->partcom);
  double tot = (double)(g->nx + 1 + 2 * g->ngx) * (double)(g->ny + 1 + 2 * g->ngy) *
               (double)(g->nz + 1 + 2 * g->ngz);
  if (tot > 4.0e9) return set_error(PXRHIP_ERR_ARG, "grid too large");
  return PXRHIP_OK;
}

static int get_table(const pxrhip_grid *g, TileTable **out) {
  int r = check_grid_desc(g);
  if (r) return r;
  if (g_tt.valid && memcmp(&g_tt.key, g, sizeof(pxrhip_grid)) == 0) {
    *out = &g_tt;
    return PXRHIP_OK;
  }
  AxisSplit ax[3];
  const int64_t nc[3] = {g->nx, g->ny, g->nz}, ntl[3] = {g->ntx, g->nty, g->ntz};
  const double x0[3] = {g->xmin, g->ymin, g->zmin}, d[3] = {g->dx, g->dy, g->dz};
  const double sd[3] = {g->sorting_dx, g->sorting_dy, g->sorting_dz};
  const double sh[3] = {g->sorting_shiftx, g->sorting_shifty, g->sorting_shiftz};
  for (int a = 0; a < 3; a++)
    if ((r = split_axis(nc[a], ntl[a], x0[a], d[a], ax[a]))) return r;
  const int nt = ax[0].nt * ax[1].nt * ax[2].nt;
  std::vector<double> sorg(3 * (size_t)nt), tbox(6 * (size_t)nt);
  std::vector<int> n3(3 * (size_t)nt);
  std::vector<int64_t> base(nt + 1);
  int64_t nb = 0, nrows = 0;
  int maxb = 0, nxmax = 1, nzmax = 1;
  for (int iz = 0; iz < ax[2].nt; iz++)
    for (int iy = 0; iy < ax[1].nt; iy++)
      for (int ix = 0; ix < ax[0].nt; ix++) {
        const int t = ix + ax[0].nt * (iy + ax[1].nt * iz);
        const int ii[3] = {ix, iy, iz};
        int64_t cells = 1;
        for (int a = 0; a < 3; a++) {
          // particle_sorting_sub: xmin2 = x_tile_min + shift, xmax2 = x_tile_max + sdx
          const double lo = ax[a].tmin[ii[a]] + sh[a];
          const double hi = ax[a].tmax[ii[a]] + sd[a];
          const double sdi = 1.0 / sd[a];  // sorting.F90:274
          const int64_t n = (int64_t)ceil((hi - lo) * sdi);
          if (n < 1 || n > (1 << 20)) return set_error(PXRHIP_ERR_ARG, "bad sorting grid");
          sorg[3 * t + a] = lo;
          tbox[6 * t + a] = ax[a].tmin[ii[a]];
          tbox[6 * t + 3 + a] = ax[a].tmax[ii[a]];
          n3[3 * t + a] = (int)n;
          cells *= n;
        }
        if (cells > (1 << 26)) return set_error(PXRHIP_ERR_ARG, "sorting grid of a tile too fine");
        base[t] = nb;
        nb += cells;
        nrows += (int64_t)n3[3 * t + 1] * n3[3 * t + 2];
        if (cells > maxb) maxb = (int)cells;
        nxmax = max(nxmax, n3[3 * t]);
        nzmax = max(nzmax, n3[3 * t + 2]);
      }
  base[nt] = nb;
  if (nb >= ((int64_t)1 << 31)) return set_error(PXRHIP_ERR_ARG, "too many sorting cells");
  free_table(g_tt);
  PXR_HIP(hipMalloc(&g_tt.sorg, sizeof(double) * sorg.size()));
  PXR_HIP(hipMalloc(&g_tt.tbox, sizeof(double) * tbox.size()));
  PXR_HIP(hipMemcpy(g_tt.tbox, tbox.data(), sizeof(double) * tbox.size(), hipMemcpyHostToDevice));
  PXR_HIP(hipMalloc(&g_tt.n3, sizeof(int) * n3.size()));
  PXR_HIP(hipMalloc(&g_tt.bin_base, sizeof(int64_t) * base.size()));
  PXR_HIP(hipMemcpy(g_tt.sorg, sorg.data(), sizeof(double) * sorg.size(), hipMemcpyHostToDevice));
  PXR_HIP(hipMemcpy(g_tt.n3, n3.data(), sizeof(int) * n3.size(), hipMemcpyHostToDevice));
  PXR_HIP(hipMemcpy(g_tt.bin_base, base.data(), sizeof(int64_t) * base.size(), hipMemcpyHostToDevice));
  g_tt.key = *g;
  g_tt.ntiles = nt;
  for (int a = 0; a < 3; a++) {
    g_tt.nt[a] = ax[a].nt;
    g_tt.n0[a] = ax[a].n0;
  }
  g_tt.nbins = nb;
  g_tt.nrows = nrows;
  g_tt.max_bins_tile = maxb;
  g_tt.nxmax = nxmax;
  g_tt.nzmax = nzmax;
  g_tt.valid = true;
  *out = &g_tt;
  return PXRHIP_OK;
}

// ---------------------------------------------------------------------------
// rebin + sort kernels
// ---------------------------------------------------------------------------
// Grid geometry of the re-binning
struct RebinGeom {
  double xmin[3], xmax[3], len[3], dx[3], sdi[3];
  int nt[3], n0[3];
  int pbound[6];            // particle boundary at x_min, x_max, y_min, ... (PXRHIP_PBOUND_*)
  const double *sorg;       // sorting-grid origin of every tile
  const int *n3;            // sorting cells of every tile
  const int64_t *bin_base;  // first bin of every tile
  const double *tbox;       // membership [tmin, tmax) of every tile
};

// Domain boundaries of one particle that left the box
// (particle_bcs_mpi_blocking, particle_boundaries.F90:1106-1300, and
// particle_bcs_tiles_and_mpi_3d, :2280-2480, which lacks the reflecting
// case: make_rebin_geom maps it to periodic there), axis by axis in x, y, z
// order as the reference examines them:
//   periodic:    x +/- L;
//   absorbing:   the particle is removed (returns false), no further axis;
//   reflecting:  x +/- dx (one cell back in) and u_d = -u_d;
//   reinjecting: x = 2 x_edge - x, u = 0, then every coordinate clamped to
//                [tmin + dx/2, tmax - dx/2] of the tile the particle is in
//                (old_tile()), and no further axis is examined.
// Each side is tested on the coordinate as it was before that axis.
template <typename OldTile>
PXR_DEV bool domain_bcs(const RebinGeom &r, double *p, double *u, bool &uchanged,
                        OldTile old_tile) {
  for (int d = 0; d < 3; d++) {
    const double v = p[d];
    for (int side = 0; side < 2; side++) {
      if (side == 0 ? !(v < r.xmin[d]) : !(v >= r.xmax[d])) continue;
      switch (r.pbound[2 * d + side]) {
        case PXRHIP_PBOUND_ABSORBING:
          return false;
        case PXRHIP_PBOUND_REFLECTING:
          p[d] = side == 0 ? v + r.dx[d] : v - r.dx[d];
          u[d] = -u[d];
          uchanged = true;
          break;
        case PXRHIP_PBOUND_REINJECTING: {
          p[d] = 2.0 * (side == 0 ? r.xmin[d] : r.xmax[d]) - p[d];
          u[0] = u[1] = u[2] = 0.0;
          uchanged = true;
          const double *tb = r.tbox + 6 * old_tile();
          for (int e = 0; e < 3; e++) {
            p[e] = fmax(p[e], tb[e] + 0.5 * r.dx[e]);
            p[e] = fmin(p[e], tb[3 + e] - 0.5 * r.dx[e]);
          }
          return true;
        }
        default:  // periodic
          p[d] = side == 0 ? v + r.len[d] : v - r.len[d];
      }
    }
  }
  return true;
}

// Sorting bin of the particle at p (inside the box after domain_bcs): tile
// index (particle_boundaries.F90:397) and its cell in that tile's sorting
// grid (sorting.F90:294-304).  ok is false (cell 0 of the tile) if the cell
// lies outside that grid.
PXR_DEV int rebin_tile(const RebinGeom &r, const double *p) {
  int it[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int t = (int)floor(((p[d] - r.xmin[d]) + r.dx[d] / 2.0) / ((double)r.n0[d] * r.dx[d]));
    it[d] = min(max(t, 0), r.nt[d] - 1);
  }
  return it[0] + r.nt[0] * (it[1] + r.nt[1] * it[2]);
}

// the tile's sorting-grid origin, cell counts and first bin
struct RebinTile {
  double so[3];
  int n3[3];
  int64_t base;
};

PXR_DEV RebinTile rebin_tile_data(const RebinGeom &r, int t) {
  RebinTile q;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    q.so[d] = r.sorg[3 * t + d];
    q.n3[d] = r.n3[3 * t + d];
  }
  q.base = r.bin_base[t];
  return q;
}

PXR_DEV int rebin_cell(const RebinGeom &r, const double *p, const RebinTile &q, bool &ok) {
  int c[3];
  ok = true;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double f = floor((p[d] - q.so[d]) * r.sdi[d]);
    if (!(f >= 0.0) || f >= (double)q.n3[d]) ok = false;  // also catches NaN
    c[d] = ok ? (int)f : 0;
  }
  const int cell = c[2] * q.n3[0] * q.n3[1] + c[1] * q.n3[0] + c[0];
  return (int)(q.base + cell);
}

struct RebinArgs {
  double *x, *y, *z;
  void *ux, *uy, *uz;       // FP64, or FP32 when mixed
  int mixed;
  const int64_t *tile_off;  // src, ntiles+1
  int ntiles;
  RebinGeom rg;
  int *key;       // global bin of every particle (-1: absorbed)
  int *hist;      // nbins counters
  unsigned long long *bad;
};

// Domain boundaries and sorting bin of every particle, counted into the bin
// histogram.  A particle that left the box is updated in place (the rare
// case; others are only read).  The particles come cell-sorted from the
// previous re-binning and few change cell in a step, so equal bins sit on
// neighbouring lanes: one atomic per run of them.  Each lane takes
// RB_ILP particles 64 apart (a wave a contiguous run of 64 * RB_ILP), and
// the phases (positions, tiles, tile tables, bins) run for all of them
// before the next, so that their chains of dependent loads overlap.
// At 256^3, 8 ppc: k_rb_keys 1.04 ms per species with 2 (four hold 134
// VGPRs, three waves per SIMD, and take 1.39 ms), 1.29 ms with 1;
// k_rb_slots 0.45 ms with 8 against 0.65 ms with 1.
constexpr int RB_ILP = 2;        // k_rb_keys
constexpr int RB_ILP_SLOTS = 8;  // k_rb_slots

__global__ __launch_bounds__(256) void k_rb_keys(RebinArgs a, i64 cap) {
  const i64 w = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const i64 i0 = w * (64 * RB_ILP) + (threadIdx.x & 63);
  const i64 ntot = a.tile_off[a.ntiles];
  const i64 n = ntot < cap ? ntot : cap;
  const RebinGeom &r = a.rg;
  double p[RB_ILP][3];
  bool keep[RB_ILP];
#pragma unroll
  for (int k = 0; k < RB_ILP; k++) {
    const i64 i = i0 + 64 * k;
    keep[k] = i < n;
    if (keep[k]) {
      p[k][0] = a.x[i];
      p[k][1] = a.y[i];
      p[k][2] = a.z[i];
    }
  }
#pragma unroll
  for (int k = 0; k < RB_ILP; k++) {
    const i64 i = i0 + 64 * k;
    double *q = p[k];
    if (!keep[k] || !(q[0] < r.xmin[0] || q[0] >= r.xmax[0] || q[1] < r.xmin[1] ||
                      q[1] >= r.xmax[1] || q[2] < r.xmin[2] || q[2] >= r.xmax[2] ||
                      q[0] != q[0] || q[1] != q[1] || q[2] != q[2]))
      continue;
    double u[3];
    if (a.mixed) {
      u[0] = ((const float *)a.ux)[i]; u[1] = ((const float *)a.uy)[i]; u[2] = ((const float *)a.uz)[i];
    } else {
      u[0] = ((const double *)a.ux)[i]; u[1] = ((const double *)a.uy)[i]; u[2] = ((const double *)a.uz)[i];
    }
    bool uch = false;
    auto old_tile = [&]() {  // the tile whose run holds particle i
      int lo = 0, hi = a.ntiles;  // last t with tile_off[t] <= i
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.tile_off[mid] <= i) lo = mid; else hi = mid;
      }
      return lo;
    };
    if (domain_bcs(r, q, u, uch, old_tile)) {
      a.x[i] = q[0];
      a.y[i] = q[1];
      a.z[i] = q[2];
      if (uch) {
        if (a.mixed) {
          ((float *)a.ux)[i] = (float)u[0]; ((float *)a.uy)[i] = (float)u[1]; ((float *)a.uz)[i] = (float)u[2];
        } else {
          ((double *)a.ux)[i] = u[0]; ((double *)a.uy)[i] = u[1]; ((double *)a.uz)[i] = u[2];
        }
      }
    } else {
      a.key[i] = -1;  // absorbed
      keep[k] = false;
    }
  }
  RebinTile tq[RB_ILP];
#pragma unroll
  for (int k = 0; k < RB_ILP; k++)
    if (keep[k]) tq[k] = rebin_tile_data(r, rebin_tile(r, p[k]));
#pragma unroll
  for (int k = 0; k < RB_ILP; k++) {
    if (!keep[k]) continue;
    bool ok;
    const int gbin = rebin_cell(r, p[k], tq[k], ok);
    if (!ok) count_bad(a.bad);
    a.key[i0 + 64 * k] = gbin;
    int head, cnt;
    wave_runs(gbin, head, cnt);
    if (cnt) atomicAdd(a.hist + gbin, cnt);
  }
}

// exclusive scan of n counts into 64-bit starts: three passes over blocks
// of SCAN_B elements
#define SCAN_B 2048
__global__ __launch_bounds__(256) void k_scan_blocks(const int *in, i64 n, i64 *bsum) {
  __shared__ i64 s[4];
  const i64 b0 = (i64)blockIdx.x * SCAN_B;
  i64 acc = 0;
  for (int k = threadIdx.x; k < SCAN_B; k += 256) {
    const i64 i = b0 + k;
    if (i < n) acc += in[i];
  }
  // the block's sum: within each wave by shuffles, then the four wave sums
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) bsum[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// in place exclusive scan of nb 64-bit values, one workgroup
__global__ __launch_bounds__(1024) void k_scan_i64_1wg(i64 *v, i64 nb, i64 *total) {
  __shared__ i64 s[16];  // inclusive sums of the sixteen waves
  __shared__ i64 carry;
  const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t == 0) carry = 0;
  __syncthreads();
  for (i64 base = 0; base < nb; base += 1024) {
    const i64 i = base + t;
    const i64 x = i < nb ? v[i] : 0;
    i64 inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const i64 y = __shfl_up(inc, off);
      if (lane >= off) inc += y;
    }
    if (lane == 63) s[wv] = inc;
    __syncthreads();
    i64 wbase = carry;
    for (int w = 0; w < wv; w++) wbase += s[w];
    if (i < nb) v[i] = wbase + inc - x;
    __syncthreads();  // s and carry read by all
    if (t == 1023) carry = wbase + inc;
    __syncthreads();
  }
  if (t == 0 && total) *total = carry;
}

// block-local exclusive scan plus the block's offset; writes out[0..n] with
// out[n] = total
// Block-local exclusive scan of SCAN_B counts plus the block's offset.  The
// counts are loaded and the starts stored with consecutive lanes on
// consecutive elements; in between they sit in LDS, where each thread scans
// SCAN_B / 256 contiguous ones (one padding slot per 16 keeps those reads off
// a single pair of banks).
__device__ __forceinline__ int scan_slot(int e) { return e + (e >> 4); }

__global__ __launch_bounds__(256) void k_scan_apply(const int *in, i64 n, const i64 *boff,
                                                    i64 *out, i64 *cursor) {
  constexpr int PER = SCAN_B / 256;
  __shared__ i64 sv[SCAN_B + SCAN_B / 16];
  __shared__ i64 s[4];
  const i64 b0 = (i64)blockIdx.x * SCAN_B;
  const int t = (int)threadIdx.x;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const i64 i = b0 + k * 256 + t;
    sv[scan_slot(k * 256 + t)] = i < n ? in[i] : 0;
  }
  __syncthreads();
  i64 acc = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) acc += sv[scan_slot(t * PER + k)];
  // inclusive scan of the 256 thread sums: within each wave by shuffles,
  // then across the four waves through LDS
  const int lane = t & 63, wv = t >> 6;
  i64 inc = acc;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const i64 y = __shfl_up(inc, off);
    if (lane >= off) inc += y;
  }
  if (lane == 63) s[wv] = inc;
  __syncthreads();
  i64 wbase = 0;
  for (int w = 0; w < wv; w++) wbase += s[w];
  i64 run = boff[blockIdx.x] + wbase + inc - acc;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int e = scan_slot(t * PER + k);
    const i64 v = sv[e];
    sv[e] = run;
    run += v;
    if (b0 + t * PER + k == n - 1) out[n] = run;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const i64 i = b0 + k * 256 + t;
    if (i < n) {
      const i64 v = sv[scan_slot(k * 256 + t)];
      out[i] = v;
      if (cursor) cursor[i] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_fill_i32_t(int *p, i64 n) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

struct ScatterArgs {
  const double *src[3];  // positions
  double *dst[3];
  const void *srcq[5];   // ux, uy, uz, gaminv, w: FP64, or FP32 when mixed
  void *dstq[5];
  int mixed;
  int npe;               // extra pid columns (FP64) carried along
  const double *srcp[PXRHIP_MAX_PID_EXTRA];
  double *dstp[PXRHIP_MAX_PID_EXTRA];
  const int64_t *src_off;
  int ntiles;
  const int *key;
  unsigned long long *cursor;  // running start of every bin
};

// Counting-sort scatter: each run of equal bins on neighbouring lanes takes
// its slots with one atomic; absorbed particles (key -1) are dropped.
__global__ __launch_bounds__(256) void k_rb_scatter(ScatterArgs a, i64 cap) {
  const i64 i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const i64 n = a.src_off[a.ntiles];
  if (i >= n || i >= cap) return;
  const int k = a.key[i];
  if (k < 0) return;
  int head, cnt;
  wave_runs(k, head, cnt);
  unsigned long long slot = 0;
  if (cnt) slot = atomicAdd(a.cursor + k, (unsigned long long)cnt);
  slot = __shfl(slot, head);
  const i64 d = (i64)slot + (__lane_id() - head);
#pragma unroll
  for (int c = 0; c < 3; c++) a.dst[c][d] = a.src[c][i];
  if (a.mixed) {
#pragma unroll
    for (int c = 0; c < 5; c++) ((float *)a.dstq[c])[d] = ((const float *)a.srcq[c])[i];
  } else {
#pragma unroll
    for (int c = 0; c < 5; c++) ((double *)a.dstq[c])[d] = ((const double *)a.srcq[c])[i];
  }
  for (int c = 0; c < a.npe; c++) a.dstp[c][d] = a.srcp[c][i];
}

// Stable counting sort, in three passes after the scan.
// 1. k_rb_slots: every run of equal bins on neighbouring lanes takes its
//    slots with one atomic and writes its source indices there, in order;
//    absorbed particles (key -1) take none.  Inside a bin the runs sit in
//    the order their atomics arrived.  A run that is not the first to
//    arrive in its bin (its slot is past the bin's start) flags the bin.
__global__ __launch_bounds__(256) void k_rb_slots(const int *key, const int64_t *src_off,
                                                  int ntiles, unsigned long long *cursor,
                                                  const i64 *bin_start, int *flag, int *perm,
                                                  i64 cap) {
  // lanes take RB_ILP_SLOTS particles 64 apart, as k_rb_keys does; the
  // atomics of all of them are issued before their slots are used
  constexpr int L = RB_ILP_SLOTS;
  const i64 w = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const i64 i0 = w * (64 * L) + (threadIdx.x & 63);
  const i64 ntot = src_off[ntiles];
  const i64 n = ntot < cap ? ntot : cap;
  int k[L], head[L];
  unsigned long long slot[L];
#pragma unroll
  for (int r = 0; r < L; r++) {
    const i64 i = i0 + 64 * r;
    k[r] = i < n ? key[i] : -1;
  }
#pragma unroll
  for (int r = 0; r < L; r++) {
    slot[r] = 0;
    head[r] = 0;
    if (k[r] < 0) continue;
    int cnt;
    wave_runs(k[r], head[r], cnt);
    if (cnt) {
      slot[r] = atomicAdd(cursor + k[r], (unsigned long long)cnt);
      if (slot[r] != (unsigned long long)bin_start[k[r]]) flag[k[r]] = 1;
    }
  }
#pragma unroll
  for (int r = 0; r < L; r++) {
    if (k[r] < 0) continue;
    const unsigned long long s0 = __shfl(slot[r], head[r]);
    perm[(i64)s0 + (__lane_id() - head[r])] = (int)(i0 + 64 * r);
  }
}

// Source indices in bin b's slots put in increasing order.  Up to 16 are
// loaded at once and sorted in registers by a fixed exchange network (no
// chain of dependent loads); longer bins take an insertion sort in memory.
PXR_DEV void order_bin(const i64 *bin_start, int b, int *perm) {
  const i64 s = bin_start[b], e = bin_start[b + 1];
  const i64 n = e - s;
  if (n < 2) return;
  if (n <= 16) {
    int v[16];
#pragma unroll
    for (int j = 0; j < 16; j++) v[j] = j < n ? perm[s + j] : INT_MAX;
#pragma unroll
    for (int i = 1; i < 16; i++)
#pragma unroll
      for (int j = i; j > 0; j--) {
        const int lo = min(v[j - 1], v[j]), hi = max(v[j - 1], v[j]);
        v[j - 1] = lo;
        v[j] = hi;
      }
#pragma unroll
    for (int j = 0; j < 16; j++)
      if (j < n) perm[s + j] = v[j];
    return;
  }
  int prev = perm[s];
  for (i64 j = s + 1; j < e; j++) {
    const int v = perm[j];
    if (v >= prev) {
      prev = v;
      continue;
    }
    i64 k = j - 1;
    while (k >= s && perm[k] > v) {
      perm[k + 1] = perm[k];
      k--;
    }
    perm[k + 1] = v;
  }
}

// 2. k_rb_order: one thread per flagged bin sorts the source indices of its
//    slots (insertion sort: the slots hold a few increasing runs), so that
//    a bin lists its particles in source order, as the reference's stable
//    counting sort does (sorting.F90:333-343), whatever order the runs
//    arrived in.  Bins filled by one run are in order already.  (Listing a
//    wave's flagged bins among 512 and sorting them one per lane took the
//    256^3 plasma's pass from 0.36 to 0.32 ms, but the bunch's from 0.7 to
//    3.2 ms: its dense flagged bins then fell to a few waves.)
__global__ __launch_bounds__(256) void k_rb_order(const i64 *bin_start, i64 nbins,
                                                  const int *flag, int *perm) {
  const i64 b = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins || !flag[b]) return;
  order_bin(bin_start, (int)b, perm);
}

struct GatherArgs {
  const double *src[3];  // positions
  double *dst[3];
  const void *srcq[5];   // ux, uy, uz, gaminv, w: FP64, or FP32 when mixed
  void *dstq[5];
  int mixed;
  int npe;               // extra pid columns (FP64) carried along
  const double *srcp[PXRHIP_MAX_PID_EXTRA];
  double *dstp[PXRHIP_MAX_PID_EXTRA];
  const int *perm;       // source index of every destination slot
  const i64 *total;      // number of particles kept
};

// 3. k_rb_gather: destination slot j takes source particle perm[j]; the
//    writes are contiguous, the reads follow the previous order.  (Two or
//    four slots per lane, their indices loaded first, made it 5% slower.)
__global__ __launch_bounds__(256) void k_rb_gather(GatherArgs a, i64 cap) {
  const i64 j = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (j >= cap || j >= *a.total) return;
  const i64 i = a.perm[j];
  // each value is read and written once here: streaming (non-temporal)
  // accesses (at 256^3, 8 ppc: 2.27 against 2.41 ms per mixed-precision
  // species, 3.14 against 3.21 ms in FP64)
#pragma unroll
  for (int c = 0; c < 3; c++) __builtin_nontemporal_store(__builtin_nontemporal_load(a.src[c] + i), a.dst[c] + j);
  if (a.mixed) {
#pragma unroll
    for (int c = 0; c < 5; c++)
      __builtin_nontemporal_store(__builtin_nontemporal_load((const float *)a.srcq[c] + i),
                                  (float *)a.dstq[c] + j);
  } else {
#pragma unroll
    for (int c = 0; c < 5; c++)
      __builtin_nontemporal_store(__builtin_nontemporal_load((const double *)a.srcq[c] + i),
                                  (double *)a.dstq[c] + j);
  }
  for (int c = 0; c < a.npe; c++)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a.srcp[c] + i), a.dstp[c] + j);
}

__global__ __launch_bounds__(256) void k_rb_tiles(int ntiles, const int64_t *bin_base,
                                                  const int64_t *bin_start, int64_t *off,
                                                  int64_t *np) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const int64_t o = bin_start[bin_base[t]];
  off[t] = o;
  if (t < ntiles) np[t] = bin_start[bin_base[t + 1]] - o;
}

// ---------------------------------------------------------------------------
// chunk table of the fused kernel
// ---------------------------------------------------------------------------
struct ChunkArgs {
  int ntiles, ch, nzmax;  // nzmax: largest number of sorting planes of a tile
  BoxGeom bg;
  i64 lds_bytes;          // LDS budget of a work item's field block
  i64 j_bytes;            // budget of its current block
  int fnode_bytes;        // LDS bytes per node of the six field arrays
  int jnode_bytes;        // LDS bytes per node of the current arrays
  const int64_t *tile_off, *tile_np, *bin_start;
  const int *n3;
  const int64_t *bin_base;
  int *count;       // chunks per unit (tile, plane)
  int *group;       // per unit: last plane of the plane group it heads, or GROUP_*
  i64 *start;       // exclusive scan of count, nunits+1 entries
  Chunk *out;
  int *nchunks, *queue;
  i64 maxchunks;    // bound on the chunks of a launch; beyond it ...
  unsigned long long *bad;  // ... chunks would be dropped, and are counted here
  int jroom;                // motion room of the current block, nodes per side
  int sk4;                  // 4 x lane stride per particle per sorting cell (0: no cap)
};

// plane roles in ChunkArgs::group (a value >= 0 is the last plane of the
// group of whole planes that this plane heads)
constexpr int GROUP_ROWS = -1;    // cut into runs of rows by walk_unit
constexpr int GROUP_MEMBER = -2;  // inside a group headed by an earlier plane
constexpr int GROUP_EMPTY = -3;   // no particles

// whether the field and current blocks of the particles of tile t in rows
// y0..y1 of planes z0..z1 fit the LDS budgets
PXR_DEV bool blocks_fit(const ChunkArgs &a, int t, int z0, int z1, int y0, int y1) {
  double ulo[3], uhi[3];
  int FR0[3], fe[3], JR0[3], je[3];
  rows_range(a.bg, t, z0, z1, y0, y1, ulo, uhi);
  blocks_for(a.bg, ulo, uhi, FR0, fe, JR0, je);
  // current block, with room for jroom nodes of motion on each side (the
  // block of k_fused is the box of the pushed particles; for a species whose
  // blocks k_fused_cut can cut down no room is needed) and for its odd
  // pitches
  const int j0 = je[0] + 2 * a.jroom, j1 = je[1] + 2 * a.jroom, j2 = je[2] + 2 * a.jroom;
  const i64 jpz = ((i64)(j0 | 1) * j1) | 1;
  return a.fnode_bytes * (i64)fe[0] * fe[1] * fe[2] <= a.lds_bytes &&
         a.jnode_bytes * jpz * j2 <= a.j_bytes;
}

// Groups of consecutive whole planes of tile t that together hold at most
// `ch` particles and whose blocks fit: sparse planes then share one work
// item instead of paying a work item's fixed cost each.  A plane too large
// to be taken whole is cut into runs of rows by walk_unit.
__global__ __launch_bounds__(256) void k_chunk_groups(ChunkArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntiles) return;
  int *grp = a.group + (i64)t * a.nzmax;
  const int nx3 = a.n3[3 * t], ny3 = a.n3[3 * t + 1], nz3 = a.n3[3 * t + 2];
  if (!a.bin_start) {
    for (int z = 0; z < a.nzmax; z++) grp[z] = GROUP_ROWS;
    return;
  }
  for (int z = nz3; z < a.nzmax; z++) grp[z] = GROUP_EMPTY;
  const i64 plane = (i64)ny3 * nx3, base = a.bin_base[t];
  auto count = [&](int z) { return a.bin_start[base + (z + 1) * plane] - a.bin_start[base + z * plane]; };
  int z = 0;
  while (z < nz3) {
    const i64 n = count(z);
    if (n <= 0) {
      grp[z++] = GROUP_EMPTY;
      continue;
    }
    if (n > a.ch || !blocks_fit(a, t, z, z, 0, ny3 - 1)) {
      grp[z++] = GROUP_ROWS;
      continue;
    }
    int z1 = z;
    i64 cnt = n;
    while (z1 + 1 < nz3) {
      const i64 m = count(z1 + 1);
      if (m <= 0 || cnt + m > a.ch || !blocks_fit(a, t, z, z1 + 1, 0, ny3 - 1)) break;
      cnt += m;
      z1++;
    }
    grp[z] = z1;
    for (int zz = z + 1; zz <= z1; zz++) grp[zz] = GROUP_MEMBER;
    z = z1 + 1;
  }
}

// A unit is one z-plane of the sorting grid of one tile.  A plane that
// heads a group of whole planes is one chunk.  Otherwise its chunks are
// runs of whole rows (cells of one y-row) of at most `ch` particles whose
// field and current blocks fit the LDS budget; rows denser than that are
// cut.  Such a chunk spans a compact block of cells: the tile's width in x,
// a few rows in y, one plane in z.  Without sorting information the tile is
// one unit cut into runs of ch particles.
template <bool WRITE>
PXR_DEV int walk_unit(const ChunkArgs &a, int t, int z, i64 pos) {
  int nch = 0;
  auto emit = [&](int64_t b, int64_t n, int zz, int zz1, int y0, int y1) {
    if (WRITE && pos + nch >= a.maxchunks) count_bad(a.bad);  // cannot happen: see the bound
    if (WRITE && pos + nch < a.maxchunks) {
      Chunk c;
      c.begin = b;
      c.n = (int)n;
      c.tile = t;
      c.z = zz;
      c.z1 = zz1;
      c.y0 = y0;
      c.y1 = y1;
      c.s = 0;
      c.pad = 0;
      if (a.sk4 > 0 && zz >= 0) {
        // lanes a few sorting cells apart: sk4 / 4 times the particles per
        // cell, to the nearest power of two (on a log scale)
        const i64 cells = (i64)a.n3[3 * t] * (y1 - y0 + 1) * (zz1 - zz + 1);
        const double want = (double)a.sk4 * (double)n / (4.0 * (double)cells);
        int p2 = 1;
        while (p2 < (1 << 20) && want > 1.4142135623730951 * p2) p2 <<= 1;
        c.s = p2;
      }
      a.out[pos + nch] = c;
    }
    nch++;
  };
  const int64_t off = a.tile_off[t], np = a.tile_np[t];
  if (np <= 0) return 0;
  if (!a.bin_start) {
    if (z != 0) return 0;
    for (int64_t b = 0; b < np; b += a.ch) emit(off + b, min((int64_t)a.ch, np - b), -1, -1, 0, 0);
    return nch;
  }
  const int nx3 = a.n3[3 * t], ny3 = a.n3[3 * t + 1], nz3 = a.n3[3 * t + 2];
  if (z >= nz3) return 0;
  const int g = a.group[(i64)t * a.nzmax + z];
  const int64_t base = a.bin_base[t] + (int64_t)z * ny3 * nx3;
  if (g >= 0) {  // whole planes z..g
    const int64_t b0 = a.bin_start[base], b1 = a.bin_start[base + (int64_t)(g - z + 1) * ny3 * nx3];
    emit(b0, b1 - b0, z, g, 0, ny3 - 1);
    return nch;
  }
  if (g != GROUP_ROWS) return 0;
  int64_t cb = 0, cnt = 0;
  int y0 = 0, ylast = 0;  // first and last row of the open chunk
  for (int y = 0; y < ny3; y++) {
    const int64_t row = base + (int64_t)y * nx3;
    const int64_t rb = a.bin_start[row], rn = a.bin_start[row + nx3] - rb;
    if (rn <= 0) continue;
    const bool fits = cnt == 0 || blocks_fit(a, t, z, z, y0, y);
    if (cnt > 0 && (cnt + rn > a.ch || !fits)) {
      emit(cb, cnt, z, z, y0, ylast);
      cnt = 0;
    }
    if (cnt == 0) {
      cb = rb;
      y0 = y;
    }
    cnt += rn;
    ylast = y;
    while (cnt > a.ch) {  // a row denser than a chunk: cut it
      emit(cb, a.ch, z, z, y0, y);
      cb += a.ch;
      cnt -= a.ch;
      y0 = y;
    }
  }
  if (cnt > 0) emit(cb, cnt, z, z, y0, ylast);
  return nch;
}

__global__ __launch_bounds__(256) void k_chunk_count(ChunkArgs a) {
  const i64 u = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < (i64)a.ntiles * a.nzmax) a.count[u] = walk_unit<false>(a, (int)(u / a.nzmax), (int)(u % a.nzmax), 0);
}

// Largest-first queue order of the work items: a counting sort of the
// items by the bit length of their particle count (32 buckets, the largest
// first), so the persistent kernel starts the long items first and ends on
// short ones.  Within a bucket the order is that of the atomics.
__global__ __launch_bounds__(256) void k_order_count(const Chunk *chunks, const int *nchunks,
                                                     int *bucket) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= *nchunks) return;
  const int n = chunks[c].n;
  atomicAdd(bucket + (n > 0 ? 31 - __clz(n) : 0), 1);
}

// bucket counts -> start slots, buckets taken from the largest down
__global__ __launch_bounds__(64) void k_order_scan(int *bucket) {
  if (threadIdx.x != 0) return;
  int s = 0;
  for (int b = 31; b >= 0; b--) {
    const int n = bucket[b];
    bucket[b] = s;
    s += n;
  }
}

__global__ __launch_bounds__(256) void k_order_scatter(const Chunk *chunks, const int *nchunks,
                                                       int *bucket, int *order) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= *nchunks) return;
  const int n = chunks[c].n;
  order[atomicAdd(bucket + (n > 0 ? 31 - __clz(n) : 0), 1)] = c;
}

__global__ __launch_bounds__(256) void k_chunk_write(ChunkArgs a) {
  const i64 nu = (i64)a.ntiles * a.nzmax;
  const i64 u = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < nu) walk_unit<true>(a, (int)(u / a.nzmax), (int)(u % a.nzmax), a.start[u]);
  if (u == 0) {
    const i64 tot = a.start[nu];
    *a.nchunks = (int)min(tot, a.maxchunks);
    *a.queue = 0;
    a.queue[1] = 0;  // overflow count (ctl[2])
  }
}

}  // namespace pxr

// ===========================================================================
// C ABI
// ===========================================================================
namespace pxr {
// dynamic LDS per workgroup when two or three workgroups (each with its
// 2 KiB of static LDS) share a CU's 160 KiB
constexpr int LDS_TWO_WG = 79616;
constexpr int LDS_THREE_WG = 52224;

// LDS that a work item of `rows` sorting rows in one plane of the widest
// tile needs: the larger of its field and current blocks, counted as
// blocks_fit counts them, for rows that start half a cell off the grid
// nodes, as tile edges do
static i64 rows_lds_bytes(const BoxGeom &bg, const TileTable &tt, int rows, int fnode,
                          int jnode) {
  const double w[3] = {tt.nxmax * bg.sd[0] * bg.di[0], rows * bg.sd[1] * bg.di[1],
                       bg.sd[2] * bg.di[2]};
  double ulo[3], uhi[3];
  for (int d = 0; d < 3; d++) {
    ulo[d] = -0.5 - 1e-7;
    uhi[d] = -0.5 + w[d] + 1e-7;
  }
  int FR0[3], fe[3], JR0[3], je[3];
  blocks_for(bg, ulo, uhi, FR0, fe, JR0, je);
  const int j0 = je[0] + 2, j1 = je[1] + 2, j2 = je[2] + 2;
  const i64 jpz = ((i64)(j0 | 1) * j1) | 1;
  const i64 fb = (i64)fnode * fe[0] * fe[1] * fe[2], jb = (i64)jnode * jpz * j2;
  return fb > jb ? fb : jb;
}
struct Tuning {
  int lds_bytes = LDS_THREE_WG;  // dynamic LDS per workgroup (with wg_per_cu > 0)
  int wg_per_cu = 0;             // persistent workgroups per CU; 0: automatic
  int spread = 1;         // lane-spread mapping in the deposit phase
  int chunk = 4096;       // default particles per work item
  // queue order of the work items: 0 as built (neighbouring items close in
  // time, which keeps shared halos in L2), 1 largest first; the latter was
  // measured slower at 256^3 (15.6 against 14.0 ms per mixed launch) and on
  // the bunch (3.66 against 2.81 ms)
  int order = 0;
  int j_lds_bytes = -1;   // cap on the LDS current block (negative: all of lds_bytes)
  int margin_pct = 0;     // motion margin of the current block, in 1/100 cell
  int jpad = 0;           // odd row/plane pitch of the LDS current block
  // largest lane stride S of the deposition (0: a sixty-fourth of the work
  // item), FP64 and mixed-precision species; -1 (the default) sets one per
  // work item (sstride_k4)
  int sstride = -1;
  int mx_sstride = -1;
  // LDS current block: 64-bit fixed point by default (3): its integer LDS
  // atomics outrun FP64 ones (at 256^3, 8 ppc the fused launch takes 14.0
  // against 19.0 ms in mixed precision, 18.6 against 22.7 ms in FP64) and
  // its sums are exact and order-independent
  int jacc = 3;           // mixed precision: 0 FP32, 1 compensated FP32, 2 FP64, 3 fixed point
  int fp64_jacc = 3;      // FP64 species: 0 FP64, 3 fixed point
  int stable_rebin = 1;   // re-binning keeps the source order inside a bin (0: arrival order)
  // launch of a mixed-precision species: its kernels take fewer registers,
  // so three workgroups with smaller LDS blocks fit on a CU
  int mx_lds_bytes = LDS_THREE_WG;
  int mx_wg_per_cu = 0;
  // motion room (nodes per side) that the work items of a mixed-precision
  // species with a fixed-point current leave around the current block of
  // their particles' old positions (k_fused_cut deposits the work items
  // whose particles outgrow it; other species keep 1): -1 automatic, i.e. 0
  // with the Esirkepov deposition (at 256^3 the launch gains 1% at 8 ppc, and
  // 19% for the 1 ppc background of the bunch workload) and 1 with the
  // direct one (order 1 at 256^3: 2% slower with 0)
  int jmargin = -1;
};
static Tuning g_tune;

// Lane stride of the deposition: lanes of a wave work on particles S apart
// inside runs of 64 S.  Measured at 256^3, 8 ppc with 8^3-node tiles
// (launch ms; 0 is a sixty-fourth of the work item):
//   order 1, direct: S 0 / 16 / 8 / 4 -> 7.05 / 5.61 / 4.94 / 4.65 mixed,
//                    8.52 / 7.55 / 6.00 / 5.52 FP64;
//   order 2:         S 0 / 16 / 8 / 4 -> 8.62 / 7.57 / 7.14 / 7.51 mixed,
//                    10.17 / 10.20 / 9.35 FP64;
//   order 3:         S 0 / 8 / 12 / 16 / 20 / 24 -> 12.63 / 13.38 / 12.68 / 12.29 /
//                    12.62 / 12.58 mixed (whole-tile work items), S 0 / 16 / 32 ->
//                    17.26 / 18.54 / 17.20 FP64 (work items of about 1,500 particles);
// the 4 ppc laser plasma chose 8 and the dense bunch 0.  The best S follows
// the particles per sorting cell of the work item: about 0.75, 1.5 and 3
// times it at orders 1, 2 and 3 (a wave's particles, and the nodes they
// deposit to, stay close while lanes sit in different cells), taken to the
// nearest power of two (17 in place of 16 at 256^3 cost 1.7%).  The
// work-item builder sets each item's S from its count and cells; returned
// here is 4 x that factor (0: no cap, FP64 at order 3).
static int sstride_k4(int order, bool mx) {
  if (order == 1) return 3;
  if (order == 2) return 6;
  return mx ? 12 : 0;
}
// the launch chosen by the last pxrhip_tiles_gather_push_deposit
static int g_last_wg_per_cu = 0, g_last_lds_bytes = 0;
}  // namespace pxr

// Charge density of a binned species, per work item: the tile blocks of
// pxrdepose_rho_on_grid_sub_openmp_3d (charge_deposition_manager.F90:516-743)
// with the nodal shapes of pxr_depose_rho_n (charge_deposition_3d.F90:
// 2081-2228).  A work item's particles add q w S_x S_y S_z / vol into an LDS
// block of rho, which is then added to rho; a particle whose stencil leaves
// the block (moved since it was binned) adds into rho directly.
struct RhoTileArgs {
  const Chunk *chunks;
  const int *nchunks;
  int *queue;
  const double *x, *y, *z;
  const void *w;  // FP64, or FP32 for a mixed-precision species
  int mixed;
  double *rho;
  GridGeom g;
  BoxGeom bg;
  double q, xmin, ymin, zmin, dxi, dyi, dzi, invvol;
  int lds_bytes;
  unsigned long long *bad;
};

template <int N>
__global__ __launch_bounds__(256) void k_rho_tiles(RhoTileArgs a) {
  extern __shared__ double srho[];
  __shared__ int s_chunk;
  const GridGeom &g = a.g;
  const int glo[3] = {-g.gx, -g.gy, -g.gz};
  const int ghi[3] = {g.nx + g.gx, g.ny + g.gy, g.nz + g.gz};
  const int nch = *a.nchunks;
  for (;;) {
    if (threadIdx.x == 0) s_chunk = atomicAdd(a.queue, 1);
    __syncthreads();
    const int c = s_chunk;
    __syncthreads();  // read by all before the next fetch overwrites it
    if (c >= nch) break;
    const Chunk ch = a.chunks[c];
    int R0[3] = {0, 0, 0}, e[3] = {0, 0, 0};
    bool blk = ch.z >= 0;
    if (blk) {
      double ulo[3], uhi[3];
      rows_range(a.bg, ch.tile, ch.z, ch.z1, ch.y0, ch.y1, ulo, uhi);
      for (int d = 0; d < 3; d++) {
        int lo, hi;
        deposit_range(a.bg, ulo[d], uhi[d], 0.0, lo, hi);
        R0[d] = max(lo, glo[d]);
        e[d] = max(min(hi, ghi[d]) - R0[d] + 1, 0);
      }
      const i64 nb64 = (i64)e[0] * e[1] * e[2];
      blk = nb64 > 0 && nb64 * (i64)sizeof(double) <= a.lds_bytes;
    }
    const int nb = blk ? e[0] * e[1] * e[2] : 0;
    for (int i = threadIdx.x; i < nb; i += 256) srho[i] = 0.0;
    __syncthreads();
    for (int k = threadIdx.x; k < ch.n; k += 256) {
      const i64 p = ch.begin + k;
      double sx[N + 1], sy[N + 1], sz[N + 1];
      const int j0 = shape_nodal<N, double>((a.x[p] - a.xmin) * a.dxi, sx);
      const int k0 = shape_nodal<N, double>((a.y[p] - a.ymin) * a.dyi, sy);
      const int l0 = shape_nodal<N, double>((a.z[p] - a.zmin) * a.dzi, sz);
      const double w = a.mixed ? (double)((const float *)a.w)[p] : ((const double *)a.w)[p];
      const double wq = a.q * w * a.invvol;
      if (blk && j0 >= R0[0] && j0 + N < R0[0] + e[0] && k0 >= R0[1] && k0 + N < R0[1] + e[1] &&
          l0 >= R0[2] && l0 + N < R0[2] + e[2]) {
        double *r = srho + (j0 - R0[0]) + e[0] * ((k0 - R0[1]) + e[1] * (l0 - R0[2]));
#pragma unroll
        for (int cc = 0; cc <= N; cc++)
#pragma unroll
          for (int bb = 0; bb <= N; bb++)
#pragma unroll
            for (int aa = 0; aa <= N; aa++)
              atomicAdd(r + aa + e[0] * (bb + e[1] * cc), sx[aa] * sy[bb] * sz[cc] * wq);
      } else if (g.inside(j0, k0, l0) && g.inside(j0 + N, k0 + N, l0 + N)) {
        double *r = a.rho + g.idx(j0, k0, l0);
#pragma unroll
        for (int cc = 0; cc <= N; cc++)
#pragma unroll
          for (int bb = 0; bb <= N; bb++)
#pragma unroll
            for (int aa = 0; aa <= N; aa++)
              unsafeAtomicAdd(r + aa + g.NX * (bb + g.NY * (i64)cc), sx[aa] * sy[bb] * sz[cc] * wq);
      } else {
        count_bad(a.bad);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += 256) {
      const double v = srho[i];
      if (v != 0.0) {
        const int ii = i % e[0], jj = (i / e[0]) % e[1], kk = i / (e[0] * e[1]);
        unsafeAtomicAdd(a.rho + g.idx(R0[0] + ii, R0[1] + jj, R0[2] + kk), v);
      }
    }
    __syncthreads();
  }
}

// Work items of a binned species (k_chunk_groups, k_chunk_count, the scan,
// k_chunk_write): runs of at most ch particles whose blocks, estimated from
// bg with fnode / jnode bytes per node, fit lds_bytes / j_bytes.  The table
// and its counters live in the library's scratch slots 8-10 and 12-15.
static int build_chunks(const TileTable *tt, const pxrhip_species *sp, const BoxGeom &bg, int ch,
                        i64 lds_bytes, i64 j_bytes, int fnode, int jnode, hipStream_t st,
                        Chunk **chunks_out, int **ctl_out, int2 **ovf_out, i64 *maxchunks_out,
                        int jroom = 1, int sk4 = 0) {
  const i64 cap = sp->capacity;
  const int nzmax = tt->nzmax;
  const i64 nunits = (i64)tt->ntiles * nzmax;
  // A plane gives one chunk if it heads a group of whole planes, none if it
  // lies inside one, else its row walk: a chunk per row that opens one
  // (<= its rows), per cut of a row denser than ch (<= cap/ch over all
  // planes) and a last one.  Hence at most nrows + cap/ch + nunits.
  const i64 maxchunks = tt->nrows + (cap / ch + 1) + nunits + 16;
  const i64 nblk = (nunits + SCAN_B - 1) / SCAN_B;
  int *cnt = (int *)scratch(8, sizeof(int) * (size_t)nunits);
  int *grp = (int *)scratch(15, sizeof(int) * (size_t)nunits);
  i64 *cstart = (i64 *)scratch(12, sizeof(i64) * (size_t)(nunits + 1));
  i64 *bsum = (i64 *)scratch(13, sizeof(i64) * (size_t)(nblk + 1));
  Chunk *chunks = (Chunk *)scratch(9, sizeof(Chunk) * (size_t)maxchunks);
  int *ctl = (int *)scratch(10, 4 * sizeof(int));
  int2 *ovf = (int2 *)scratch(14, sizeof(int2) * (size_t)maxchunks);
  if (!cnt || !grp || !cstart || !bsum || !chunks || !ctl || !ovf) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
  ChunkArgs ca;
  ca.ntiles = tt->ntiles;
  ca.ch = ch;
  ca.nzmax = nzmax;
  ca.bg = bg;
  ca.lds_bytes = lds_bytes;
  ca.j_bytes = j_bytes;
  ca.fnode_bytes = fnode;
  ca.jnode_bytes = jnode;
  ca.tile_off = sp->tile_off;
  ca.tile_np = sp->tile_np;
  ca.bin_start = sp->bin_start;
  ca.n3 = tt->n3;
  ca.bin_base = tt->bin_base;
  ca.count = cnt;
  ca.group = grp;
  ca.start = cstart;
  ca.out = chunks;
  ca.nchunks = ctl;
  ca.queue = ctl + 1;
  ca.maxchunks = maxchunks;
  ca.bad = rt().d_bad;
  ca.jroom = jroom;
  ca.sk4 = sk4;
  hipLaunchKernelGGL(k_chunk_groups, dim3(grid1d(tt->ntiles, 64)), dim3(64), 0, st, ca);
  hipLaunchKernelGGL(k_chunk_count, dim3(grid1d(nunits, 256)), dim3(256), 0, st, ca);
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nblk), dim3(256), 0, st, cnt, nunits, bsum);
  hipLaunchKernelGGL(k_scan_i64_1wg, dim3(1), dim3(1024), 0, st, bsum, nblk, (i64 *)nullptr);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nblk), dim3(256), 0, st, cnt, nunits, bsum,
                     cstart, (i64 *)nullptr);
  hipLaunchKernelGGL(k_chunk_write, dim3(grid1d(nunits, 256)), dim3(256), 0, st, ca);
  PXR_LAUNCHED();
  *chunks_out = chunks;
  *ctl_out = ctl;
  *ovf_out = ovf;
  *maxchunks_out = maxchunks;
  return PXRHIP_OK;
}

extern "C" {

int pxrhip_tiles_nbins(const pxrhip_grid *grid, int64_t *nbins) {
  PXR_INIT();
  TileTable *tt;
  int r = get_table(grid, &tt);
  if (r) return r;
  if (nbins) *nbins = tt->nbins;
  return PXRHIP_OK;
}

int pxrhip_set_tuning(const char *key, int64_t value) {
  if (!key) return set_error(PXRHIP_ERR_ARG, "null tuning key");
  if (!strcmp(key, "fused.lds_bytes")) {
    if (value < 4096 || value > 160 * 1024 - 1024)
      return set_error(PXRHIP_ERR_ARG, "fused.lds_bytes must be in [4096, 159744]");
    g_tune.lds_bytes = (int)(value & ~7);
  } else if (!strcmp(key, "fused.wg_per_cu")) {
    if (value < 0 || value > 16)
      return set_error(PXRHIP_ERR_ARG, "fused.wg_per_cu must be in [0, 16] (0: automatic)");
    g_tune.wg_per_cu = (int)value;
  } else if (!strcmp(key, "fused.spread")) {
    g_tune.spread = value != 0;
  } else if (!strcmp(key, "fused.j_lds_bytes")) {
    g_tune.j_lds_bytes = value < 0 ? -1 : (value > (1 << 30) ? (1 << 30) : (int)value);
  } else if (!strcmp(key, "mixed.lds_bytes")) {
    if (value < 4096 || value > 160 * 1024 - 1024)
      return set_error(PXRHIP_ERR_ARG, "mixed.lds_bytes must be in [4096, 159744]");
    g_tune.mx_lds_bytes = (int)(value & ~7);
  } else if (!strcmp(key, "mixed.wg_per_cu")) {
    if (value < 0 || value > 16)
      return set_error(PXRHIP_ERR_ARG, "mixed.wg_per_cu must be in [0, 16] (0: automatic)");
    g_tune.mx_wg_per_cu = (int)value;
  } else if (!strcmp(key, "mixed.jacc")) {
    if (value < 0 || value > 3)
      return set_error(PXRHIP_ERR_ARG, "mixed.jacc must be 0, 1, 2 or 3");
    g_tune.jacc = (int)value;
  } else if (!strcmp(key, "fused.jacc")) {
    if (value != 0 && value != 3) return set_error(PXRHIP_ERR_ARG, "fused.jacc must be 0 or 3");
    g_tune.fp64_jacc = (int)value;
  } else if (!strcmp(key, "fused.sstride")) {
    if (value < -1 || value > (1 << 20)) return set_error(PXRHIP_ERR_ARG, "fused.sstride must be in [-1, 2^20]");
    g_tune.sstride = (int)value;
  } else if (!strcmp(key, "mixed.sstride")) {
    if (value < -1 || value > (1 << 20)) return set_error(PXRHIP_ERR_ARG, "mixed.sstride must be in [-1, 2^20]");
    g_tune.mx_sstride = (int)value;
  } else if (!strcmp(key, "fused.jpad")) {
    g_tune.jpad = value != 0;
  } else if (!strcmp(key, "rebin.stable")) {
    g_tune.stable_rebin = value != 0;
  } else if (!strcmp(key, "fused.order")) {
    g_tune.order = value != 0;
  } else if (!strcmp(key, "leaf.bin_min")) {
    if (value < 0) return set_error(PXRHIP_ERR_ARG, "leaf.bin_min must be >= 0 (0: never)");
    rt().leaf_bin_min = value;
  } else if (!strcmp(key, "leaf.bin_chunk")) {
    if (value < 1 || value > (1 << 20)) return set_error(PXRHIP_ERR_ARG, "leaf.bin_chunk must be in [1, 2^20]");
    rt().leaf_bin_chunk = (int)value;
  } else if (!strcmp(key, "fused.margin_pct")) {
    if (value < 0 || value > 200) return set_error(PXRHIP_ERR_ARG, "fused.margin_pct must be in [0, 200]");
    g_tune.margin_pct = (int)value;
  } else if (!strcmp(key, "fused.chunk")) {
    if (value < 1 || value > (1 << 20)) return set_error(PXRHIP_ERR_ARG, "fused.chunk must be in [1, 2^20]");
    g_tune.chunk = (int)value;
  } else if (!strcmp(key, "mixed.jmargin")) {
    if (value < -1 || value > 1) return set_error(PXRHIP_ERR_ARG, "mixed.jmargin must be -1 (automatic), 0 or 1");
    g_tune.jmargin = (int)value;
  } else {
    return set_error(PXRHIP_ERR_ARG, "unknown tuning key '%s'", key);
  }
  return PXRHIP_OK;
}

int pxrhip_get_tuning(const char *key, int64_t *value) {
  if (!key || !value) return set_error(PXRHIP_ERR_ARG, "null argument");
  if (!strcmp(key, "fused.lds_bytes")) *value = g_tune.lds_bytes;
  else if (!strcmp(key, "fused.wg_per_cu")) *value = g_tune.wg_per_cu;
  else if (!strcmp(key, "fused.spread")) *value = g_tune.spread;
  else if (!strcmp(key, "fused.chunk")) *value = g_tune.chunk;
  else if (!strcmp(key, "mixed.jmargin")) *value = g_tune.jmargin;
  else if (!strcmp(key, "fused.j_lds_bytes")) *value = g_tune.j_lds_bytes;
  else if (!strcmp(key, "fused.margin_pct")) *value = g_tune.margin_pct;
  else if (!strcmp(key, "fused.jpad")) *value = g_tune.jpad;
  else if (!strcmp(key, "rebin.stable")) *value = g_tune.stable_rebin;
  else if (!strcmp(key, "fused.order")) *value = g_tune.order;
  else if (!strcmp(key, "leaf.bin_min")) *value = rt().leaf_bin_min;
  else if (!strcmp(key, "leaf.bin_chunk")) *value = rt().leaf_bin_chunk;
  else if (!strcmp(key, "fused.sstride")) *value = g_tune.sstride;
  else if (!strcmp(key, "mixed.sstride")) *value = g_tune.mx_sstride;
  else if (!strcmp(key, "mixed.jacc")) *value = g_tune.jacc;
  else if (!strcmp(key, "fused.jacc")) *value = g_tune.fp64_jacc;
  else if (!strcmp(key, "mixed.lds_bytes")) *value = g_tune.mx_lds_bytes;
  else if (!strcmp(key, "mixed.wg_per_cu")) *value = g_tune.mx_wg_per_cu;
  else if (!strcmp(key, "last.wg_per_cu")) *value = g_last_wg_per_cu;
  else if (!strcmp(key, "device.num_cus")) *value = rt().device >= 0 ? rt().num_cus : 0;
  else if (!strcmp(key, "last.lds_bytes")) *value = g_last_lds_bytes;
  else return set_error(PXRHIP_ERR_ARG, "unknown tuning key '%s'", key);
  return PXRHIP_OK;
}

int pxrhip_tiles_stats(int64_t *nitems, int64_t *noverflow) {
  PXR_INIT();
  int *ctl = (int *)scratch(10, 4 * sizeof(int));
  if (!ctl) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
  int h[4] = {0, 0, 0, 0};
  PXR_HIP(hipMemcpyAsync(h, ctl, sizeof(h), hipMemcpyDeviceToHost, rt().stream));
  PXR_HIP(hipStreamSynchronize(rt().stream));
  if (nitems) *nitems = h[0];
  if (noverflow) *noverflow = h[2];
  return PXRHIP_OK;
}

int pxrhip_tiles_check(void) {
  PXR_INIT();
  unsigned long long n = 0;
  int r = take_bad(&n);
  if (r) return r;
  if (n)
    return set_error(PXRHIP_ERR_BOUNDS, "%llu particle(s) were outside the grid arrays or sorting "
                     "grids and were skipped", n);
  return PXRHIP_OK;
}

static int check_species(const pxrhip_species *s, bool need_bins) {
  if (!s) return set_error(PXRHIP_ERR_ARG, "null species");
  if (!s->x || !s->y || !s->z || !s->ux || !s->uy || !s->uz || !s->gaminv || !s->w)
    return set_error(PXRHIP_ERR_ARG, "null particle array");
  if (!s->tile_off || !s->tile_np) return set_error(PXRHIP_ERR_ARG, "null tile offsets");
  if (need_bins && !s->bin_start) return set_error(PXRHIP_ERR_ARG, "null bin_start");
  if (s->capacity < 0 || s->capacity > ((int64_t)1 << 31) - 1)
    return set_error(PXRHIP_ERR_ARG, "capacity must be in [0, 2^31)");
  if (s->precision != 0 && s->precision != 1)
    return set_error(PXRHIP_ERR_ARG, "precision must be 0 (FP64) or 1 (mixed)");
  if (s->npid_extra < 0 || s->npid_extra > PXRHIP_MAX_PID_EXTRA)
    return set_error(PXRHIP_ERR_ARG, "npid_extra must be in [0, %d]", PXRHIP_MAX_PID_EXTRA);
  for (int c = 0; c < s->npid_extra; c++)
    if (!s->pid_extra[c]) return set_error(PXRHIP_ERR_ARG, "null pid column %d", c + 2);
  return PXRHIP_OK;
}

static RebinGeom make_rebin_geom(const pxrhip_grid *grid, const TileTable *tt) {
  RebinGeom rg;
  const double xm[3] = {grid->xmin, grid->ymin, grid->zmin}, d[3] = {grid->dx, grid->dy, grid->dz};
  const double sd[3] = {grid->sorting_dx, grid->sorting_dy, grid->sorting_dz};
  const int64_t nc[3] = {grid->nx, grid->ny, grid->nz};
  for (int k = 0; k < 3; k++) {
    rg.xmin[k] = xm[k];
    rg.len[k] = (double)nc[k] * d[k];
    rg.xmax[k] = xm[k] + rg.len[k];
    rg.dx[k] = d[k];
    rg.sdi[k] = 1.0 / sd[k];
    rg.nt[k] = tt->nt[k];
    rg.n0[k] = tt->n0[k];
  }
  // particle_bcs_tiles_and_mpi_3d (partcom 0) has no reflecting case: its
  // SELECT CASE falls through to the periodic wrap (particle_boundaries.F90:
  // 2308-2333 and the five faces after it)
  for (int k = 0; k < 6; k++) {
    const int b = (int)grid->pbound[k];
    rg.pbound[k] = (grid->partcom == 0 && b == PXRHIP_PBOUND_REFLECTING) ? PXRHIP_PBOUND_PERIODIC : b;
  }
  rg.sorg = tt->sorg;
  rg.n3 = tt->n3;
  rg.bin_base = tt->bin_base;
  rg.tbox = tt->tbox;
  return rg;
}

int pxrhip_tiles_depose_rho(const pxrhip_grid *grid, const pxrhip_species *sp, double *rho,
                            int64_t order) {
  PXR_INIT();
  TileTable *tt;
  int r = get_table(grid, &tt);
  if (r) return r;
  if ((r = check_species(sp, true))) return r;
  if (order < 1 || order > 3) return set_error(PXRHIP_ERR_UNSUPPORTED, "shape order must be 1..3");
  if (!rho) return set_error(PXRHIP_ERR_ARG, "null rho");
  hipStream_t st = rt().stream;
  BoxGeom bg;
  const double x0[3] = {grid->xmin, grid->ymin, grid->zmin};
  const double dd[3] = {grid->dx, grid->dy, grid->dz};
  const double sd[3] = {grid->sorting_dx, grid->sorting_dy, grid->sorting_dz};
  const int64_t nn[3] = {grid->nx, grid->ny, grid->nz}, ng[3] = {grid->ngx, grid->ngy, grid->ngz};
  for (int d = 0; d < 3; d++) {
    bg.x0[d] = x0[d];
    bg.di[d] = 1.0 / dd[d];
    bg.sd[d] = sd[d];
    bg.mv[d] = 0.0;
    bg.glo[d] = (int)-ng[d];
    bg.ghi[d] = (int)(nn[d] + ng[d]);
  }
  bg.s = 0.5;
  bg.N = bg.M = (int)order;
  bg.depo = 3;  // the nodal index ranges of the direct deposition cover those of rho
  bg.sorg = tt->sorg;
  bg.tbox = tt->tbox;
  bg.n3 = tt->n3;
  // four workgroups per CU, each with a quarter of the LDS less its static part
  const i64 lds_bytes = 38912;
  Chunk *chunks;
  int *ctl;
  int2 *ovf;
  i64 maxchunks;
  if ((r = build_chunks(tt, sp, bg, g_tune.chunk, lds_bytes, lds_bytes, 0, 8, st, &chunks, &ctl,
                        &ovf, &maxchunks)))
    return r;
  RhoTileArgs a;
  a.chunks = chunks;
  a.nchunks = ctl;
  a.queue = ctl + 1;
  a.x = sp->x; a.y = sp->y; a.z = sp->z;
  a.w = sp->w;
  a.mixed = sp->precision == 1;
  a.rho = rho;
  GridGeom g;
  g.nx = (int)grid->nx; g.ny = (int)grid->ny; g.nz = (int)grid->nz;
  g.gx = (int)grid->ngx; g.gy = (int)grid->ngy; g.gz = (int)grid->ngz;
  g.NX = grid->nx + 1 + 2 * grid->ngx;
  g.NY = grid->ny + 1 + 2 * grid->ngy;
  a.g = g;
  a.bg = bg;
  a.q = sp->q;
  a.xmin = grid->xmin; a.ymin = grid->ymin; a.zmin = grid->zmin;
  a.dxi = 1.0 / grid->dx; a.dyi = 1.0 / grid->dy; a.dzi = 1.0 / grid->dz;
  a.invvol = 1.0 / (grid->dx * grid->dy * grid->dz);
  a.lds_bytes = (int)lds_bytes;
  a.bad = rt().d_bad;
  const dim3 gr((unsigned)(rt().num_cus * 4)), bl(256);
  if (order == 1) hipLaunchKernelGGL(k_rho_tiles<1>, gr, bl, (size_t)lds_bytes, st, a);
  else if (order == 2) hipLaunchKernelGGL(k_rho_tiles<2>, gr, bl, (size_t)lds_bytes, st, a);
  else hipLaunchKernelGGL(k_rho_tiles<3>, gr, bl, (size_t)lds_bytes, st, a);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_tiles_rebin_sort(const pxrhip_grid *grid, pxrhip_species *src, pxrhip_species *dst) {
  PXR_INIT();
  TileTable *tt;
  int r = get_table(grid, &tt);
  if (r) return r;
  if ((r = check_species(src, false)) || (r = check_species(dst, true))) return r;
  if (dst->precision != src->precision)
    return set_error(PXRHIP_ERR_ARG, "source and destination species differ in precision");
  if (dst->npid_extra != src->npid_extra)
    return set_error(PXRHIP_ERR_ARG, "source and destination species differ in pid columns");
  if (dst->capacity < src->capacity)
    return set_error(PXRHIP_ERR_CAPACITY, "dst capacity smaller than src capacity");
  const i64 cap = src->capacity;
  const i64 nbins = tt->nbins;
  hipStream_t st = rt().stream;
  int *key = (int *)scratch(4, sizeof(int) * (size_t)(cap > 1 ? cap : 1));
  int *hist = (int *)scratch(5, sizeof(int) * (size_t)nbins);
  const i64 nblk = (nbins + SCAN_B - 1) / SCAN_B;
  i64 *bsum = (i64 *)scratch(6, sizeof(i64) * (size_t)(nblk + 1));
  unsigned long long *cursor = (unsigned long long *)scratch(7, sizeof(i64) * (size_t)nbins);
  if (!key || !hist || !bsum || !cursor) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
  const RebinGeom rg = make_rebin_geom(grid, tt);
  RebinArgs a;
  a.x = src->x;
  a.y = src->y;
  a.z = src->z;
  a.ux = src->ux;
  a.uy = src->uy;
  a.uz = src->uz;
  a.mixed = src->precision == 1;
  a.tile_off = src->tile_off;
  a.ntiles = tt->ntiles;
  a.rg = rg;
  a.key = key;
  a.hist = hist;
  a.bad = rt().d_bad;
  hipLaunchKernelGGL(k_fill_i32_t, dim3(grid1d(nbins, 256)), dim3(256), 0, st, hist, nbins);
  hipLaunchKernelGGL(k_rb_keys, dim3(grid1d((cap + RB_ILP - 1) / RB_ILP, 256)), dim3(256), 0, st, a,
                     cap);
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nblk), dim3(256), 0, st, hist, nbins, bsum);
  hipLaunchKernelGGL(k_scan_i64_1wg, dim3(1), dim3(1024), 0, st, bsum, nblk, (i64 *)nullptr);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nblk), dim3(256), 0, st, hist, nbins, bsum,
                     (i64 *)dst->bin_start, (i64 *)cursor);
  ScatterArgs sa;
  double *S[3] = {src->x, src->y, src->z}, *D[3] = {dst->x, dst->y, dst->z};
  void *SQ[5] = {src->ux, src->uy, src->uz, src->gaminv, src->w};
  void *DQ[5] = {dst->ux, dst->uy, dst->uz, dst->gaminv, dst->w};
  for (int k = 0; k < 3; k++) {
    sa.src[k] = S[k];
    sa.dst[k] = D[k];
  }
  for (int k = 0; k < 5; k++) {
    sa.srcq[k] = SQ[k];
    sa.dstq[k] = DQ[k];
  }
  sa.mixed = src->precision == 1;
  sa.npe = (int)src->npid_extra;
  for (int c = 0; c < PXRHIP_MAX_PID_EXTRA; c++) {
    sa.srcp[c] = c < sa.npe ? src->pid_extra[c] : nullptr;
    sa.dstp[c] = c < sa.npe ? dst->pid_extra[c] : nullptr;
  }
  sa.src_off = src->tile_off;
  sa.ntiles = tt->ntiles;
  sa.key = key;
  sa.cursor = cursor;
  if (g_tune.stable_rebin) {
    int *perm = (int *)scratch(16, sizeof(int) * (size_t)(cap > 1 ? cap : 1));
    int *flag = (int *)scratch(18, sizeof(int) * (size_t)nbins);  // bins to order
    if (!perm || !flag) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
    hipLaunchKernelGGL(k_fill_i32_t, dim3(grid1d(nbins, 256)), dim3(256), 0, st, flag, nbins);
    hipLaunchKernelGGL(k_rb_slots, dim3(grid1d((cap + RB_ILP_SLOTS - 1) / RB_ILP_SLOTS, 256)),
                       dim3(256), 0, st,
                       (const int *)key,
                       (const int64_t *)src->tile_off, tt->ntiles, cursor,
                       (const i64 *)dst->bin_start, flag, perm, cap);
    hipLaunchKernelGGL(k_rb_order, dim3(grid1d(nbins, 256)), dim3(256), 0, st,
                       (const i64 *)dst->bin_start, nbins, flag, perm);
    GatherArgs ga;
    for (int k = 0; k < 3; k++) {
      ga.src[k] = S[k];
      ga.dst[k] = D[k];
    }
    for (int k = 0; k < 5; k++) {
      ga.srcq[k] = SQ[k];
      ga.dstq[k] = DQ[k];
    }
    ga.mixed = sa.mixed;
    ga.npe = sa.npe;
    for (int c = 0; c < PXRHIP_MAX_PID_EXTRA; c++) {
      ga.srcp[c] = sa.srcp[c];
      ga.dstp[c] = sa.dstp[c];
    }
    ga.perm = perm;
    ga.total = (const i64 *)dst->bin_start + nbins;
    hipLaunchKernelGGL(k_rb_gather, dim3(grid1d(cap, 256)), dim3(256), 0, st, ga, cap);
  } else {
    hipLaunchKernelGGL(k_rb_scatter, dim3(grid1d(cap, 256)), dim3(256), 0, st, sa, cap);
  }
  hipLaunchKernelGGL(k_rb_tiles, dim3(grid1d(tt->ntiles + 1, 256)), dim3(256), 0, st, tt->ntiles,
                     tt->bin_base, dst->bin_start, dst->tile_off, dst->tile_np);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

}  // extern "C"

// The tiled gather + push + deposit of one species (k_fused), restricted
// to the phases in `phase` (PHASE_PUSH, PHASE_DEPOSIT or both).
static int tiles_fused(const pxrhip_grid *grid, const pxrhip_species *sp, const double *exg,
                       const double *eyg, const double *ezg, const double *bxg, const double *byg,
                       const double *bzg, double *jx, double *jy, double *jz, double dt,
                       int64_t order, int64_t l_lower_order_in_v, int64_t pusher,
                       int64_t depo_algo, int64_t chunk, int phase) {
  PXR_INIT();
  TileTable *tt;
  int r = get_table(grid, &tt);
  if (r) return r;
  if ((r = check_species(sp, phase == PHASE_DEPOSIT))) return r;
  if (order < 1 || order > 3) return set_error(PXRHIP_ERR_UNSUPPORTED, "shape order must be 1..3");
  if (pusher != 0 && pusher != 1) return set_error(PXRHIP_ERR_ARG, "pusher must be 0 (Boris) or 1 (Vay)");
  if (depo_algo != 0 && depo_algo != 3) return set_error(PXRHIP_ERR_ARG, "depo_algo must be 0 or 3");
  if ((phase & PHASE_PUSH) && (!exg || !eyg || !ezg || !bxg || !byg || !bzg))
    return set_error(PXRHIP_ERR_ARG, "null field array");
  if ((phase & PHASE_DEPOSIT) && (!jx || !jy || !jz)) return set_error(PXRHIP_ERR_ARG, "null current array");
  if (!(dt > 0)) return set_error(PXRHIP_ERR_ARG, "dt must be > 0");
  const int ch = chunk > 0 ? (int)(chunk < (1 << 20) ? chunk : (1 << 20)) : g_tune.chunk;
  hipStream_t st = rt().stream;
  // block geometry of the work items
  BoxGeom bg;
  {
    const double x0[3] = {grid->xmin, grid->ymin, grid->zmin};
    const double dd[3] = {grid->dx, grid->dy, grid->dz};
    const double sd[3] = {grid->sorting_dx, grid->sorting_dy, grid->sorting_dz};
    const int64_t nn[3] = {grid->nx, grid->ny, grid->nz}, ng[3] = {grid->ngx, grid->ngy, grid->ngz};
    const double margin = g_tune.margin_pct / 100.0;
    for (int d = 0; d < 3; d++) {
      bg.x0[d] = x0[d];
      bg.di[d] = 1.0 / dd[d];
      bg.sd[d] = sd[d];
      const double cfl = dt / dd[d];  // |v| < c = 1: at most dt/dx cells per step
      bg.mv[d] = margin < cfl ? margin : cfl;
      bg.glo[d] = (int)-ng[d];
      bg.ghi[d] = (int)(nn[d] + ng[d]);
    }
    bg.s = 0.5;
    bg.N = (int)order;
    bg.M = l_lower_order_in_v ? (int)order - 1 : (int)order;
    bg.depo = depo_algo ? 3 : 0;
    bg.sorg = tt->sorg;
    bg.tbox = tt->tbox;
    bg.n3 = tt->n3;
  }
  const bool mx = sp->precision == 1;
  const int esize = mx ? 4 : 8;
  // LDS current: mixed.jacc (0 FP32, 1 compensated FP32, 2 FP64, 3 fixed
  // point) for a mixed-precision species, fused.jacc (0 FP64, 3 fixed point)
  // for an FP64 one
  const int jacc = mx ? g_tune.jacc : g_tune.fp64_jacc;
  const int jnode = (mx && jacc == 0) ? 12 : 24;  // 3 FP32 | 3 x 8 bytes
  i64 lds_bytes = mx ? g_tune.mx_lds_bytes : g_tune.lds_bytes;
  int wg_per_cu = mx ? g_tune.mx_wg_per_cu : g_tune.wg_per_cu;
  if (wg_per_cu == 0) {
    // automatic: three workgroups per CU when a work item of four sorting
    // rows of the widest tile fits their LDS, else two with more each
    const bool three = rows_lds_bytes(bg, *tt, 4, 6 * esize, jnode) <= LDS_THREE_WG;
    wg_per_cu = three ? 3 : 2;
    lds_bytes = three ? LDS_THREE_WG : LDS_TWO_WG;
  }
  if (!mx && order == 3 && depo_algo && wg_per_cu > 2) {
    // that kernel holds registers for two waves per SIMD: two workgroups
    // per CU, each with the LDS of two
    wg_per_cu = 2;
    lds_bytes = LDS_TWO_WG;
  }
  const i64 j_bytes = g_tune.j_lds_bytes < 0 || g_tune.j_lds_bytes > lds_bytes ? lds_bytes
                                                                              : g_tune.j_lds_bytes;
  Chunk *chunks;
  int *ctl;
  int2 *ovf;
  i64 maxchunks;
  // blocks that the work items' phases stage: a phase left out needs no LDS
  const int fnode = (phase & PHASE_PUSH) ? 6 * esize : 0;
  const int jnode_b = (phase & PHASE_DEPOSIT) ? jnode : 0;
  // a mixed-precision species with a fixed-point current: work items whose
  // current outgrows its LDS block are deposited by k_fused_cut
  const bool defer = mx && jacc == 3 && (phase & PHASE_DEPOSIT);
  // lane stride of the deposition: a tuning value, else one per work item
  const int sstride_set = mx ? g_tune.mx_sstride : g_tune.sstride;
  const int sk4 = sstride_set >= 0 ? 0 : sstride_k4((int)order, mx);
  if ((r = build_chunks(tt, sp, bg, ch, lds_bytes, j_bytes, fnode, jnode_b, st, &chunks, &ctl,
                        &ovf, &maxchunks,
                        !defer ? 1 : g_tune.jmargin >= 0 ? g_tune.jmargin : (depo_algo ? 1 : 0),
                        sk4)))
    return r;
  // fused kernel
  FusedArgs a;
  a.chunks = chunks;
  a.nchunks = ctl;
  a.queue = ctl + 1;
  a.x = sp->x; a.y = sp->y; a.z = sp->z;
  a.ux = a.uy = a.uz = a.gi = nullptr;
  a.w = nullptr;
  a.fux = a.fuy = a.fuz = a.fgi = nullptr;
  a.fw = nullptr;
  if (mx) {
    a.fux = (float *)sp->ux; a.fuy = (float *)sp->uy; a.fuz = (float *)sp->uz;
    a.fgi = (float *)sp->gaminv; a.fw = (const float *)sp->w;
  } else {
    a.ux = sp->ux; a.uy = sp->uy; a.uz = sp->uz; a.gi = sp->gaminv; a.w = sp->w;
  }
  const double *F[6] = {exg, eyg, ezg, bxg, byg, bzg};
  for (int f = 0; f < 6; f++) a.F[f] = F[f];
  a.J[0] = jx; a.J[1] = jy; a.J[2] = jz;
  GridGeom g;
  g.nx = (int)grid->nx; g.ny = (int)grid->ny; g.nz = (int)grid->nz;
  g.gx = (int)grid->ngx; g.gy = (int)grid->ngy; g.gz = (int)grid->ngz;
  g.NX = grid->nx + 1 + 2 * grid->ngx;
  g.NY = grid->ny + 1 + 2 * grid->ngy;
  a.g = g;
  a.xmin = grid->xmin; a.ymin = grid->ymin; a.zmin = grid->zmin;
  a.dxi = 1.0 / grid->dx; a.dyi = 1.0 / grid->dy; a.dzi = 1.0 / grid->dz;
  a.s = 0.5;
  a.cst = pusher == 1 ? sp->q * dt / sp->m : sp->q * dt * 0.5 / sp->m;
  a.dt = dt;
  a.ec = make_esirk_const<double>(sp->q, grid->xmin, grid->ymin, grid->zmin, dt, grid->dx,
                                  grid->dy, grid->dz);
  a.q = sp->q; a.dx = grid->dx; a.dy = grid->dy; a.dz = grid->dz;
  a.bad = rt().d_bad;
  a.bg = bg;
  a.lds_bytes = (int)lds_bytes;
  a.j_bytes = (int)j_bytes;
  a.spread = g_tune.spread;
  a.jpad = g_tune.jpad;
  a.sstride = sstride_set > 0 ? sstride_set : 0;  // else per work item (Chunk::s)
  a.jacc = jacc;
  a.ovf_count = ctl + 2;
  a.ovf = ovf;
  a.ovf_cap = maxchunks < INT_MAX ? (int)maxchunks : INT_MAX;
  a.order = nullptr;
  if (g_tune.order) {
    int *bucket = (int *)scratch(20, sizeof(int) * 32);
    int *order = (int *)scratch(21, sizeof(int) * (size_t)maxchunks);
    if (!bucket || !order) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
    hipLaunchKernelGGL(k_fill_i32_t, dim3(1), dim3(256), 0, st, bucket, (i64)32);
    hipLaunchKernelGGL(k_order_count, dim3(grid1d(maxchunks, 256)), dim3(256), 0, st,
                       (const Chunk *)chunks, (const int *)ctl, bucket);
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(64), 0, st, bucket);
    hipLaunchKernelGGL(k_order_scatter, dim3(grid1d(maxchunks, 256)), dim3(256), 0, st,
                       (const Chunk *)chunks, (const int *)ctl, bucket, order);
    PXR_LAUNCHED();
    a.order = order;
  }
  a.phase = phase | (defer ? PHASE_CUT : 0);
  // largest |current contribution| per unit of fx_weight (|q w| times the
  // displacement in cells, or times |v| for the direct form): Esirkepov's
  // running sums of shape differences along an axis are at most the
  // displacement along it and its transverse weights at most 3/2
  // (esirkepov_3d.F90:3427-3450); direct shapes are at most 1
  a.fxb = depo_algo ? 1.0 / (grid->dx * grid->dy * grid->dz)
                    : 1.5 * fmax(a.ec.invdtdx, fmax(a.ec.invdtdy, a.ec.invdtdz));
  const int nwg = rt().num_cus * wg_per_cu;
  g_last_wg_per_cu = wg_per_cu;
  g_last_lds_bytes = (int)lds_bytes;
  const size_t lds = (size_t)lds_bytes;
  const int key = (int)(order * 100 + (l_lower_order_in_v ? 10 : 0) + pusher * 2 + (depo_algo ? 1 : 0));
  if (!mx)
    return jacc == 3 ? launch_fused_fp64_fx(key, a, nwg, lds, st)
                     : launch_fused_fp64(key, a, nwg, lds, st);
  if (jacc == 3) return launch_fused_mixed_fx(key, a, nwg, lds, st);
  return jacc == 2 ? launch_fused_mixed_f64acc(key, a, nwg, lds, st)
                   : launch_fused_mixed_f32(key, a, nwg, lds, st);
}

extern "C" {

int pxrhip_tiles_gather_push_deposit(const pxrhip_grid *grid, const pxrhip_species *sp,
                                     const double *exg, const double *eyg, const double *ezg,
                                     const double *bxg, const double *byg, const double *bzg,
                                     double *jx, double *jy, double *jz, double dt, int64_t order,
                                     int64_t l_lower_order_in_v, int64_t pusher,
                                     int64_t depo_algo, int64_t chunk) {
  return tiles_fused(grid, sp, exg, eyg, ezg, bxg, byg, bzg, jx, jy, jz, dt, order,
                     l_lower_order_in_v, pusher, depo_algo, chunk, PHASE_PUSH | PHASE_DEPOSIT);
}

int pxrhip_tiles_gather_push(const pxrhip_grid *grid, const pxrhip_species *sp,
                             const double *exg, const double *eyg, const double *ezg,
                             const double *bxg, const double *byg, const double *bzg, double dt,
                             int64_t order, int64_t l_lower_order_in_v, int64_t pusher,
                             int64_t chunk) {
  return tiles_fused(grid, sp, exg, eyg, ezg, bxg, byg, bzg, nullptr, nullptr, nullptr, dt,
                     order, l_lower_order_in_v, pusher, 0, chunk, PHASE_PUSH);
}

int pxrhip_tiles_depose_currents(const pxrhip_grid *grid, const pxrhip_species *sp, double *jx,
                                 double *jy, double *jz, double dt, int64_t order,
                                 int64_t depo_algo, int64_t chunk) {
  return tiles_fused(grid, sp, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, jx, jy, jz,
                     dt, order, 1, 0, depo_algo, chunk, PHASE_DEPOSIT);
}

}  // extern "C"
