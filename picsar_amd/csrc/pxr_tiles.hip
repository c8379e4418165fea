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
  if (t.sorg || t.n3 || t.bin_base || t.tbox) (void)hipStreamSynchronize(rt().stream);
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
struct RebinArgs {
  double *x, *y, *z;
  const int64_t *tile_off;  // src, ntiles+1
  int ntiles;
  double xmin[3], xmax[3], len[3], dx[3], sdi[3];
  int nt[3], n0[3];
  const double *sorg;
  const int *n3;
  const int64_t *bin_base;
  int *key;       // global bin of every particle
  int *hist;      // nbins counters
  unsigned long long *bad;
};

// periodic wrap (particle_bcs_mpi_blocking, periodic case), tile index
// (particle_boundaries.F90:397) and sorting key (sorting.F90:294-304)
__global__ __launch_bounds__(256) void k_rb_keys(RebinArgs a, i64 cap) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const i64 n = a.tile_off[a.ntiles];
  const bool act = i < n && i < cap;
  int gbin = -1;
  if (act) {
    double p[3] = {a.x[i], a.y[i], a.z[i]};
    int it[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double v = p[d];
      if (v < a.xmin[d]) p[d] = v + a.len[d];
      if (v >= a.xmax[d]) p[d] = v - a.len[d];
      int t = (int)floor(((p[d] - a.xmin[d]) + a.dx[d] / 2.0) / ((double)a.n0[d] * a.dx[d]));
      it[d] = min(max(t, 0), a.nt[d] - 1);
    }
    a.x[i] = p[0];
    a.y[i] = p[1];
    a.z[i] = p[2];
    const int t = it[0] + a.nt[0] * (it[1] + a.nt[1] * it[2]);
    int c[3];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double f = floor((p[d] - a.sorg[3 * t + d]) * a.sdi[d]);
      const int n3 = a.n3[3 * t + d];
      if (!(f >= 0.0) || f >= (double)n3) ok = false;  // also catches NaN
      c[d] = ok ? (int)f : 0;
    }
    if (!ok) count_bad(a.bad);
    const int cell = c[2] * a.n3[3 * t] * a.n3[3 * t + 1] + c[1] * a.n3[3 * t] + c[0];
    gbin = (int)(a.bin_base[t] + cell);
    a.key[i] = gbin;
  }
  int rank, cnt, leader;
  wave_peers(act, gbin, rank, cnt, leader);
  if (act && rank == 0) atomicAdd(a.hist + gbin, cnt);
}

// exclusive scan of n counts into 64-bit starts: three passes over blocks
// of SCAN_B elements
#define SCAN_B 2048
__global__ __launch_bounds__(256) void k_scan_blocks(const int *in, i64 n, i64 *bsum) {
  __shared__ i64 s[256];
  const i64 b0 = (i64)blockIdx.x * SCAN_B;
  i64 acc = 0;
  for (int k = threadIdx.x; k < SCAN_B; k += 256) {
    const i64 i = b0 + k;
    if (i < n) acc += in[i];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}

// in place exclusive scan of nb 64-bit values, one workgroup
__global__ __launch_bounds__(1024) void k_scan_i64_1wg(i64 *v, i64 nb, i64 *total) {
  __shared__ i64 s[1024];
  __shared__ i64 carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (i64 base = 0; base < nb; base += 1024) {
    const i64 i = base + threadIdx.x;
    const i64 x = i < nb ? v[i] : 0;
    s[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const i64 y = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += y;
      __syncthreads();
    }
    const i64 incl = s[threadIdx.x];
    if (i < nb) v[i] = carry + incl - x;
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

// block-local exclusive scan plus the block's offset; writes out[0..n] with
// out[n] = total
__global__ __launch_bounds__(256) void k_scan_apply(const int *in, i64 n, const i64 *boff,
                                                    i64 *out, i64 *cursor) {
  __shared__ i64 s[256];
  const i64 b0 = (i64)blockIdx.x * SCAN_B;
  constexpr int PER = SCAN_B / 256;
  i64 loc[PER];
  i64 acc = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const i64 i = b0 + (i64)threadIdx.x * PER + k;
    loc[k] = i < n ? in[i] : 0;
    acc += loc[k];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const i64 y = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += y;
    __syncthreads();
  }
  i64 run = boff[blockIdx.x] + s[threadIdx.x] - acc;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const i64 i = b0 + (i64)threadIdx.x * PER + k;
    if (i < n) {
      out[i] = run;
      if (cursor) cursor[i] = run;
    }
    run += loc[k];
    if (i == n - 1) out[n] = run;
  }
}

__global__ __launch_bounds__(256) void k_fill_i32_t(int *p, i64 n) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

struct ScatterArgs {
  const double *src[8];
  double *dst[8];
  const int64_t *src_off;
  int ntiles;
  const int *key;
  unsigned long long *cursor;  // running start of every bin
};

__global__ __launch_bounds__(256) void k_rb_scatter(ScatterArgs a, i64 cap) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const i64 n = a.src_off[a.ntiles];
  const bool act = i < n && i < cap;
  const int k = act ? a.key[i] : -1;
  int rank, cnt, leader;
  wave_peers(act, k, rank, cnt, leader);
  unsigned long long slot = 0;
  if (act && rank == 0) slot = atomicAdd(a.cursor + k, (unsigned long long)cnt);
  slot = __shfl(slot, leader);
  if (!act) return;
  const i64 d = (i64)slot + rank;
#pragma unroll
  for (int c = 0; c < 8; c++) a.dst[c][d] = a.src[c][i];
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

PXR_DEV void gather_range(const BoxGeom &b, double ulo, double uhi, int &lo, int &hi) {
  const int N = b.N, M = b.M;
  auto jn = [&](double v) { return (N == 2 ? f_nint(v) : f_floor(v)) - (N == 1 ? 0 : 1); };
  auto js = [&](double v) {
    return (M == 0 || M == 2) ? f_floor((v + 0.5) - b.s) - (M == 2 ? 1 : 0)
                              : f_floor(v - b.s) - (M == 3 ? 1 : 0);
  };
  lo = min(jn(ulo), js(ulo));
  hi = max(jn(uhi) + N, js(uhi) + M);
}

PXR_DEV void deposit_range(const BoxGeom &b, double ulo, double uhi, double m, int &lo, int &hi) {
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
// y0..y1 of plane z (intersected with the tile's membership)
PXR_DEV void rows_range(const BoxGeom &b, int t, int z, int y0, int y1, double *ulo, double *uhi) {
  const double *so = b.sorg + 3 * t, *tb = b.tbox + 6 * t;
  const int c0[3] = {0, y0, z};
  const int c1[3] = {b.n3[3 * t] - 1, y1, z};
  for (int d = 0; d < 3; d++) {
    const double plo = fmax(so[d] + b.sd[d] * c0[d], tb[d]);
    const double phi = fmin(so[d] + b.sd[d] * (c1[d] + 1), tb[3 + d]);
    ulo[d] = (plo - b.x0[d]) * b.di[d] - 1e-7;
    uhi[d] = (phi - b.x0[d]) * b.di[d] + 1e-7;
  }
}

// field block (FR0, fe) and current block (JR0, je) for that range, clamped
// to the arrays; returns the doubles they take in LDS (6 + 3 arrays)
PXR_DEV i64 blocks_for(const BoxGeom &b, const double *ulo, const double *uhi, int *FR0, int *fe,
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
  int z, y0, y1;  // sorting-grid plane and rows of the particles; z < 0: unknown
  int pad;
};

struct ChunkArgs {
  int ntiles, ch, nzmax;  // nzmax: largest number of sorting planes of a tile
  BoxGeom bg;
  i64 lds_doubles;        // LDS budget of a work item's field + current blocks
  i64 j_doubles;          // budget of its current block
  const int64_t *tile_off, *tile_np, *bin_start;
  const int *n3;
  const int64_t *bin_base;
  int *count;       // chunks per unit (tile, plane)
  i64 *start;       // exclusive scan of count, nunits+1 entries
  Chunk *out;
  int *nchunks, *queue;
  i64 maxchunks;
};

// A unit is one z-plane of the sorting grid of one tile.  Its chunks are
// runs of whole rows (cells of one y-row) of at most `ch` particles whose
// field and current blocks fit the LDS budget; rows denser than that are
// cut.  A chunk therefore spans a compact block of
// cells: the tile's width in x, a few rows in y, one plane in z.  Without
// sorting information the tile is one unit cut into runs of ch particles.
template <bool WRITE>
PXR_DEV int walk_unit(const ChunkArgs &a, int t, int z, i64 pos) {
  int nch = 0;
  auto emit = [&](int64_t b, int64_t n, int zz, int y0, int y1) {
    if (WRITE && pos + nch < a.maxchunks) {
      Chunk c;
      c.begin = b;
      c.n = (int)n;
      c.tile = t;
      c.z = zz;
      c.y0 = y0;
      c.y1 = y1;
      c.pad = 0;
      a.out[pos + nch] = c;
    }
    nch++;
  };
  const int64_t off = a.tile_off[t], np = a.tile_np[t];
  if (np <= 0) return 0;
  if (!a.bin_start) {
    if (z != 0) return 0;
    for (int64_t b = 0; b < np; b += a.ch) emit(off + b, min((int64_t)a.ch, np - b), -1, 0, 0);
    return nch;
  }
  const int nx3 = a.n3[3 * t], ny3 = a.n3[3 * t + 1], nz3 = a.n3[3 * t + 2];
  if (z >= nz3) return 0;
  const int64_t base = a.bin_base[t] + (int64_t)z * ny3 * nx3;
  int64_t cb = 0, cnt = 0;
  int y0 = 0, ylast = 0;  // first and last row of the open chunk
  for (int y = 0; y < ny3; y++) {
    const int64_t row = base + (int64_t)y * nx3;
    const int64_t rb = a.bin_start[row], rn = a.bin_start[row + nx3] - rb;
    if (rn <= 0) continue;
    bool fits = true;
    if (cnt > 0) {
      double ulo[3], uhi[3];
      int FR0[3], fe[3], JR0[3], je[3];
      rows_range(a.bg, t, z, y0, y, ulo, uhi);
      blocks_for(a.bg, ulo, uhi, FR0, fe, JR0, je);
      // current block, with room for the odd pitches of k_fused
      const i64 jpz = ((i64)(je[0] | 1) * je[1]) | 1;
      fits = 6 * (i64)fe[0] * fe[1] * fe[2] <= a.lds_doubles && 3 * jpz * je[2] <= a.j_doubles;
    }
    if (cnt > 0 && (cnt + rn > a.ch || !fits)) {
      emit(cb, cnt, z, y0, ylast);
      cnt = 0;
    }
    if (cnt == 0) {
      cb = rb;
      y0 = y;
    }
    cnt += rn;
    ylast = y;
    while (cnt > a.ch) {  // a row denser than a chunk: cut it
      emit(cb, a.ch, z, y0, y);
      cb += a.ch;
      cnt -= a.ch;
      y0 = y;
    }
  }
  if (cnt > 0) emit(cb, cnt, z, y0, ylast);
  return nch;
}

__global__ __launch_bounds__(256) void k_chunk_count(ChunkArgs a) {
  const i64 u = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < (i64)a.ntiles * a.nzmax) a.count[u] = walk_unit<false>(a, (int)(u / a.nzmax), (int)(u % a.nzmax), 0);
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

// ---------------------------------------------------------------------------
// fused gather + push + deposit
// ---------------------------------------------------------------------------
struct FusedArgs {
  const Chunk *chunks;
  const int *nchunks;
  int *queue;
  double *x, *y, *z, *ux, *uy, *uz, *gi;
  const double *w;
  const double *F[6];
  double *J[3];
  GridGeom g;
  double xmin, ymin, zmin, dxi, dyi, dzi, s;
  double cst;        // Boris: q*dt/(2m); Vay: q*dt/m
  double dt;
  EsirkConst<double> ec;
  double q, dx, dy, dz;  // direct deposition
  unsigned long long *bad;
  BoxGeom bg;            // work-item block geometry
  int lds_doubles;       // dynamic LDS size in doubles (field and current blocks)
  int j_doubles;         // largest current block kept in LDS, in doubles
  int spread;            // lanes take particles S apart
  int jpad;              // odd pitches for the LDS current block
  int *ovf_count;        // work items whose blocks do not fit in LDS ...
  int2 *ovf;             // ... (work item, stage): 0 untouched, 1 pushed
  int ovf_cap;
};

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

struct LdsAcc {
  double *p[3];
  int i0, j0, k0, sy, sz;
  PXR_DEV void add(int c, int i, int j, int k, double v) const {
    atomicAdd(p[c] + ((i - i0) + sy * (j - j0) + sz * (k - k0)), v);
  }
};

struct GAcc {
  double *p[3];
  GridGeom g;
  PXR_DEV void add(int c, int i, int j, int k, double v) const {
    unsafeAtomicAdd(p[c] + g.idx(i, j, k), v);
  }
};

PXR_DEV bool in_block(const int *lo, const int *hi, const int *R0, const int *e) {
  return lo[0] >= R0[0] && lo[1] >= R0[1] && lo[2] >= R0[2] && hi[0] < R0[0] + e[0] &&
         hi[1] < R0[1] + e[1] && hi[2] < R0[2] + e[2];
}

// Gather, push and position push of particle p.  Fields are read from the
// LDS block when the stencil lies inside it (views fl), else from global
// memory (fg).  Extends the box of the current the particle will deposit.
template <int N, bool LO, int PUSHER, int DEPO, typename VL, typename VG>
PXR_DEV void gpd_push(const FusedArgs &a, i64 p, const VL *fl, const VG *fg, const int *FR0,
                      const int *fe, Box6 &dep) {
  constexpr int M = LO ? N - 1 : N;
  const GridGeom &g = a.g;
  const double x = a.x[p], y = a.y[p], z = a.z[p];
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
  double u[3] = {a.ux[p], a.uy[p], a.uz[p]};
  double gi = a.gi[p];
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
  int b0[3], b1[3];
  if (DEPO == 0) {
    EsirkepovIn<double> in;
    esirkepov_prepare<N, double>(in, xn, yn, zn, u[0], u[1], u[2], gi, a.w[p], a.ec);
    if (!esirkepov_box<N, double>(in, b0, b1)) return;
  } else {
    DirectIn<double> in;
    direct_prepare<N, double>(in, xn, yn, zn, u[0], u[1], u[2], gi, a.w[p], a.q, a.xmin, a.ymin,
                              a.zmin, a.dt, a.dx, a.dy, a.dz, a.s);
    direct_box<N>(in, b0, b1);
  }
  for (int d = 0; d < 3; d++) {
    dep.lo[d] = min(dep.lo[d], b0[d]);
    dep.hi[d] = max(dep.hi[d], b1[d]);
  }
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

// Persistent kernel over the work items.  Per work item:
//   1. the node box of its field block, from the sorting-grid cells that
//      hold its particles (or, without sorting information, from one pass
//      over the positions); the six field blocks are staged into LDS;
//   2. gather (from LDS; a particle whose stencil leaves the block, e.g.
//      after a push without re-sorting, reads global memory), push and
//      position push, lanes on consecutive particles; the box of the
//      currents is reduced across the workgroup;
//   3. the current block is zeroed in LDS and the pushed particles are
//      deposited with LDS atomics; lanes of a wave take particles S apart
//      (one or more cells apart in a cell-sorted run) so that their atomics
//      rarely meet;
//   4. the current block is added to global J.
// A work item whose field block does not fit in LDS is left whole to
// k_fused_overflow; one whose current block does not fit is left to it
// after the push.
template <int BS, int N, bool LO, int PUSHER, int DEPO>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2))) void k_fused(FusedArgs a) {
  constexpr int M = LO ? N - 1 : N;
  extern __shared__ double smem[];
  __shared__ int s_chunk;
  __shared__ int s_box[6];
  const GridGeom &g = a.g;
  const int glo[3] = {-g.gx, -g.gy, -g.gz};
  const int ghi[3] = {g.nx + g.gx, g.ny + g.gy, g.nz + g.gz};
  const double x0[3] = {a.xmin, a.ymin, a.zmin}, di[3] = {a.dxi, a.dyi, a.dzi};
  const int nch = *a.nchunks;
  for (;;) {
    if (threadIdx.x == 0) {
      s_chunk = atomicAdd(a.queue, 1);
      for (int d = 0; d < 3; d++) {
        s_box[d] = INT_MAX;
        s_box[3 + d] = INT_MIN;
      }
    }
    __syncthreads();
    const int c = s_chunk;
    if (c >= nch) break;
    const Chunk ch = a.chunks[c];
    const i64 p0 = ch.begin;
    // ---- 1. field block -------------------------------------------------
    double ulo[3], uhi[3];
    if (ch.z >= 0) {
      rows_range(a.bg, ch.tile, ch.z, ch.y0, ch.y1, ulo, uhi);
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
    if (nb64 <= 0 || 6 * nb64 > a.lds_doubles) {
      if (threadIdx.x == 0) {
        const int slot = atomicAdd(a.ovf_count, 1);
        if (slot < a.ovf_cap) a.ovf[slot] = make_int2(c, 0);
      }
      __syncthreads();
      continue;
    }
    const int nb = (int)nb64;
    for (int f = 0; f < 6; f++) {
      const double *src = a.F[f];
      for (int idx = threadIdx.x; idx < nb; idx += BS) {
        const int i = FR0[0] + idx % fe[0];
        const int j = FR0[1] + (idx / fe[0]) % fe[1];
        const int k = FR0[2] + idx / (fe[0] * fe[1]);
        smem[f * nb + idx] = src[g.idx(i, j, k)];
      }
    }
    __syncthreads();
    // ---- 2. gather, push, position push ---------------------------------
    Box6 dep;
    box_reset(dep);
    {
      LdsView<double> fl[6];
      FieldView<double> fg[6];
      for (int f = 0; f < 6; f++) {
        fl[f] = LdsView<double>{smem + f * nb, FR0[0], FR0[1], FR0[2], fe[0], fe[0] * fe[1]};
        fg[f] = FieldView<double>{a.F[f] + g.idx(0, 0, 0), g.NX, g.NX * g.NY};
      }
      for (int k = threadIdx.x; k < ch.n; k += BS)
        gpd_push<N, LO, PUSHER, DEPO>(a, p0 + k, fl, fg, FR0, fe, dep);
    }
    wave_minmax(dep);
    __syncthreads();  // field block no longer read
    if ((threadIdx.x & 63) == 0)
      for (int d = 0; d < 3; d++) {
        atomicMin(&s_box[d], dep.lo[d]);
        atomicMax(&s_box[3 + d], dep.hi[d]);
      }
    __syncthreads();
    int R0[3], e[3];
    for (int d = 0; d < 3; d++) {
      R0[d] = max(s_box[d], glo[d]);
      const int R1 = min(s_box[3 + d], ghi[d]);
      e[d] = R1 >= R0[d] ? R1 - R0[d] + 1 : 0;
    }
    // LDS layout of the current block: odd row and plane pitches, so that
    // lanes on cells an even number apart fall on different banks
    const int py = a.jpad ? (e[0] | 1) : e[0];
    const int pzr = py * e[1];
    const int pz = a.jpad ? (pzr | 1) : pzr;
    const i64 nj64 = (i64)pz * e[2];
    if ((i64)e[0] * e[1] * e[2] <= 0 || 3 * nj64 > a.j_doubles) {
      // pushed; the deposition is left to the overflow kernel
      if (threadIdx.x == 0 && (i64)e[0] * e[1] * e[2] > 0) {
        const int slot = atomicAdd(a.ovf_count, 1);
        if (slot < a.ovf_cap) a.ovf[slot] = make_int2(c, 1);
      }
      __syncthreads();
      continue;
    }
    const int nj = (int)nj64;
    for (int idx = threadIdx.x; idx < 3 * nj; idx += BS) smem[idx] = 0.0;
    __syncthreads();
    // ---- 3. deposit -------------------------------------------------------
    {
      const LdsAcc la{{smem, smem + nj, smem + 2 * nj}, R0[0], R0[1], R0[2], py, pz};
      const int S = (ch.n + 63) >> 6;
      for (int kk = threadIdx.x; kk < S * 64; kk += BS) {
        const int k = a.spread ? (kk & 63) * S + (kk >> 6) : kk;
        if (k < ch.n && !gpd_deposit<N, DEPO>(a, p0 + k, la)) count_bad(a.bad);
      }
    }
    __syncthreads();
    // ---- 4. flush ----------------------------------------------------------
    const int ne = e[0] * e[1] * e[2];
    for (int idx = threadIdx.x; idx < ne; idx += BS) {
      const int i = idx % e[0];
      const int j = (idx / e[0]) % e[1];
      const int k = idx / (e[0] * e[1]);
      const i64 gi_ = g.idx(R0[0] + i, R0[1] + j, R0[2] + k);
      const int li = i + py * j + pz * k;
      for (int c3 = 0; c3 < 3; c3++) {
        const double v = smem[c3 * nj + li];
        if (v != 0.0) unsafeAtomicAdd(a.J[c3] + gi_, v);
      }
    }
    __syncthreads();
  }
}

// Work items whose blocks did not fit in LDS: gather from global memory and
// deposit with global atomics, one workgroup per work item.  Stage 1: the
// particles were already pushed by k_fused.
template <int BS, int N, bool LO, int PUSHER, int DEPO>
__global__ __launch_bounds__(BS) void k_fused_overflow(FusedArgs a) {
  const int n = min(*a.ovf_count, a.ovf_cap);
  const GridGeom &g = a.g;
  FieldView<double> fg[6];
  for (int f = 0; f < 6; f++) fg[f] = FieldView<double>{a.F[f] + g.idx(0, 0, 0), g.NX, g.NX * g.NY};
  const GAcc ga{{a.J[0], a.J[1], a.J[2]}, g};
  const int none[3] = {0, 0, 0};  // empty block: every gather reads global memory
  for (int o = blockIdx.x; o < n; o += gridDim.x) {
    const int2 ov = a.ovf[o];
    const Chunk ch = a.chunks[ov.x];
    if (ov.y == 0) {
      Box6 dep;
      box_reset(dep);
      for (int k = threadIdx.x; k < ch.n; k += BS)
        gpd_push<N, LO, PUSHER, DEPO>(a, ch.begin + k, fg, fg, none, none, dep);
    }
    __syncthreads();
    const int S = (ch.n + 63) >> 6;
    for (int kk = threadIdx.x; kk < S * 64; kk += BS) {
      const int k = a.spread ? (kk & 63) * S + (kk >> 6) : kk;
      if (k < ch.n && !gpd_deposit<N, DEPO>(a, ch.begin + k, ga)) count_bad(a.bad);
    }
    __syncthreads();
  }
}

}  // namespace pxr

// ===========================================================================
// C ABI
// ===========================================================================
static const int FUSED_BS = 256;

namespace pxr {
struct Tuning {
  int lds_bytes = 79616;  // dynamic LDS per workgroup: two workgroups (with their 2 KiB
                          // of static LDS) fit in a CU's 160 KiB
  int wg_per_cu = 2;      // persistent workgroups per CU
  int spread = 1;         // lane-spread mapping in the deposit phase
  int chunk = 4096;       // default particles per work item
  int j_lds_bytes = -1;   // cap on the LDS current block (negative: all of lds_bytes)
  int margin_pct = 0;     // motion margin of the current block, in 1/100 cell
  int jpad = 0;           // odd row/plane pitch of the LDS current block
};
static Tuning g_tune;
}  // namespace pxr

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
    if (value < 1 || value > 16) return set_error(PXRHIP_ERR_ARG, "fused.wg_per_cu must be in [1, 16]");
    g_tune.wg_per_cu = (int)value;
  } else if (!strcmp(key, "fused.spread")) {
    g_tune.spread = value != 0;
  } else if (!strcmp(key, "fused.j_lds_bytes")) {
    g_tune.j_lds_bytes = value < 0 ? -1 : (value > (1 << 30) ? (1 << 30) : (int)value);
  } else if (!strcmp(key, "fused.jpad")) {
    g_tune.jpad = value != 0;
  } else if (!strcmp(key, "fused.margin_pct")) {
    if (value < 0 || value > 200) return set_error(PXRHIP_ERR_ARG, "fused.margin_pct must be in [0, 200]");
    g_tune.margin_pct = (int)value;
  } else if (!strcmp(key, "fused.chunk")) {
    if (value < 1 || value > (1 << 20)) return set_error(PXRHIP_ERR_ARG, "fused.chunk must be in [1, 2^20]");
    g_tune.chunk = (int)value;
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
  else if (!strcmp(key, "fused.j_lds_bytes")) *value = g_tune.j_lds_bytes;
  else if (!strcmp(key, "fused.margin_pct")) *value = g_tune.margin_pct;
  else if (!strcmp(key, "fused.jpad")) *value = g_tune.jpad;
  else return set_error(PXRHIP_ERR_ARG, "unknown tuning key '%s'", key);
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
  return PXRHIP_OK;
}

int pxrhip_tiles_rebin_sort(const pxrhip_grid *grid, pxrhip_species *src, pxrhip_species *dst) {
  PXR_INIT();
  TileTable *tt;
  int r = get_table(grid, &tt);
  if (r) return r;
  if ((r = check_species(src, false)) || (r = check_species(dst, true))) return r;
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
  RebinArgs a;
  a.x = src->x;
  a.y = src->y;
  a.z = src->z;
  a.tile_off = src->tile_off;
  a.ntiles = tt->ntiles;
  const double xm[3] = {grid->xmin, grid->ymin, grid->zmin}, d[3] = {grid->dx, grid->dy, grid->dz};
  const double sd[3] = {grid->sorting_dx, grid->sorting_dy, grid->sorting_dz};
  const int64_t nc[3] = {grid->nx, grid->ny, grid->nz};
  for (int k = 0; k < 3; k++) {
    a.xmin[k] = xm[k];
    a.len[k] = (double)nc[k] * d[k];
    a.xmax[k] = xm[k] + a.len[k];
    a.dx[k] = d[k];
    a.sdi[k] = 1.0 / sd[k];
    a.nt[k] = tt->nt[k];
    a.n0[k] = tt->n0[k];
  }
  a.sorg = tt->sorg;
  a.n3 = tt->n3;
  a.bin_base = tt->bin_base;
  a.key = key;
  a.hist = hist;
  a.bad = rt().d_bad;
  hipLaunchKernelGGL(k_fill_i32_t, dim3(grid1d(nbins, 256)), dim3(256), 0, st, hist, nbins);
  hipLaunchKernelGGL(k_rb_keys, dim3(grid1d(cap, 256)), dim3(256), 0, st, a, cap);
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nblk), dim3(256), 0, st, hist, nbins, bsum);
  hipLaunchKernelGGL(k_scan_i64_1wg, dim3(1), dim3(1024), 0, st, bsum, nblk, (i64 *)nullptr);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nblk), dim3(256), 0, st, hist, nbins, bsum,
                     (i64 *)dst->bin_start, (i64 *)cursor);
  ScatterArgs sa;
  double *S[8] = {src->x, src->y, src->z, src->ux, src->uy, src->uz, src->gaminv, src->w};
  double *D[8] = {dst->x, dst->y, dst->z, dst->ux, dst->uy, dst->uz, dst->gaminv, dst->w};
  for (int k = 0; k < 8; k++) {
    sa.src[k] = S[k];
    sa.dst[k] = D[k];
  }
  sa.src_off = src->tile_off;
  sa.ntiles = tt->ntiles;
  sa.key = key;
  sa.cursor = cursor;
  hipLaunchKernelGGL(k_rb_scatter, dim3(grid1d(cap, 256)), dim3(256), 0, st, sa, cap);
  hipLaunchKernelGGL(k_rb_tiles, dim3(grid1d(tt->ntiles + 1, 256)), dim3(256), 0, st, tt->ntiles,
                     tt->bin_base, dst->bin_start, dst->tile_off, dst->tile_np);
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

int pxrhip_tiles_gather_push_deposit(const pxrhip_grid *grid, const pxrhip_species *sp,
                                     const double *exg, const double *eyg, const double *ezg,
                                     const double *bxg, const double *byg, const double *bzg,
                                     double *jx, double *jy, double *jz, double dt, int64_t order,
                                     int64_t l_lower_order_in_v, int64_t pusher,
                                     int64_t depo_algo, int64_t chunk) {
  PXR_INIT();
  TileTable *tt;
  int r = get_table(grid, &tt);
  if (r) return r;
  if ((r = check_species(sp, false))) return r;
  if (order < 1 || order > 3) return set_error(PXRHIP_ERR_UNSUPPORTED, "shape order must be 1..3");
  if (pusher != 0 && pusher != 1) return set_error(PXRHIP_ERR_ARG, "pusher must be 0 (Boris) or 1 (Vay)");
  if (depo_algo != 0 && depo_algo != 3) return set_error(PXRHIP_ERR_ARG, "depo_algo must be 0 or 3");
  if (!exg || !eyg || !ezg || !bxg || !byg || !bzg || !jx || !jy || !jz)
    return set_error(PXRHIP_ERR_ARG, "null field array");
  const int ch = chunk > 0 ? (int)(chunk < (1 << 20) ? chunk : (1 << 20)) : g_tune.chunk;
  hipStream_t st = rt().stream;
  // chunk table
  const i64 cap = sp->capacity;
  const int nzmax = tt->nzmax;
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
  const i64 lds_doubles = g_tune.lds_bytes / 8;
  const i64 j_doubles = g_tune.j_lds_bytes < 0 || g_tune.j_lds_bytes / 8 > lds_doubles
                            ? lds_doubles : g_tune.j_lds_bytes / 8;
  const i64 nunits = (i64)tt->ntiles * nzmax;
  const i64 maxchunks = tt->nrows + (cap / ch + 1) + nunits + 16;
  const i64 nblk = (nunits + SCAN_B - 1) / SCAN_B;
  int *cnt = (int *)scratch(8, sizeof(int) * (size_t)nunits);
  i64 *cstart = (i64 *)scratch(12, sizeof(i64) * (size_t)(nunits + 1));
  i64 *bsum = (i64 *)scratch(13, sizeof(i64) * (size_t)(nblk + 1));
  Chunk *chunks = (Chunk *)scratch(9, sizeof(Chunk) * (size_t)maxchunks);
  int *ctl = (int *)scratch(10, 4 * sizeof(int));
  int2 *ovf = (int2 *)scratch(14, sizeof(int2) * (size_t)maxchunks);
  if (!cnt || !cstart || !bsum || !chunks || !ctl || !ovf) return set_error(PXRHIP_ERR_HIP, "scratch allocation failed");
  ChunkArgs ca;
  ca.ntiles = tt->ntiles;
  ca.ch = ch;
  ca.nzmax = nzmax;
  ca.bg = bg;
  ca.lds_doubles = lds_doubles;
  ca.j_doubles = j_doubles;
  ca.tile_off = sp->tile_off;
  ca.tile_np = sp->tile_np;
  ca.bin_start = sp->bin_start;
  ca.n3 = tt->n3;
  ca.bin_base = tt->bin_base;
  ca.count = cnt;
  ca.start = cstart;
  ca.out = chunks;
  ca.nchunks = ctl;
  ca.queue = ctl + 1;
  ca.maxchunks = maxchunks;
  hipLaunchKernelGGL(k_chunk_count, dim3(grid1d(nunits, 256)), dim3(256), 0, st, ca);
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nblk), dim3(256), 0, st, cnt, nunits, bsum);
  hipLaunchKernelGGL(k_scan_i64_1wg, dim3(1), dim3(1024), 0, st, bsum, nblk, (i64 *)nullptr);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nblk), dim3(256), 0, st, cnt, nunits, bsum,
                     cstart, (i64 *)nullptr);
  hipLaunchKernelGGL(k_chunk_write, dim3(grid1d(nunits, 256)), dim3(256), 0, st, ca);
  PXR_LAUNCHED();
  // fused kernel
  FusedArgs a;
  a.chunks = chunks;
  a.nchunks = ctl;
  a.queue = ctl + 1;
  a.x = sp->x; a.y = sp->y; a.z = sp->z;
  a.ux = sp->ux; a.uy = sp->uy; a.uz = sp->uz; a.gi = sp->gaminv; a.w = sp->w;
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
  a.lds_doubles = (int)lds_doubles;
  a.j_doubles = (int)j_doubles;
  a.spread = g_tune.spread;
  a.jpad = g_tune.jpad;
  a.ovf_count = ctl + 2;
  a.ovf = ovf;
  a.ovf_cap = maxchunks < INT_MAX ? (int)maxchunks : INT_MAX;
  const int nwg = rt().num_cus * g_tune.wg_per_cu;
  dim3 grd(nwg), blk(FUSED_BS);
  const size_t lds = (size_t)g_tune.lds_bytes;
  const int key = (int)(order * 100 + (l_lower_order_in_v ? 10 : 0) + pusher * 2 + (depo_algo ? 1 : 0));
#define PXR_FUSED_CASE(N_, LO_, P_, D_)                                                          \
  case N_ * 100 + (LO_ ? 10 : 0) + P_ * 2 + (D_ ? 1 : 0):                                        \
    if (lds > 65536)                                                                             \
      PXR_HIP(hipFuncSetAttribute((const void *)&k_fused<FUSED_BS, N_, LO_, P_, D_>,             \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));       \
    hipLaunchKernelGGL((k_fused<FUSED_BS, N_, LO_, P_, D_>), grd, blk, lds, st, a);              \
    hipLaunchKernelGGL((k_fused_overflow<FUSED_BS, N_, LO_, P_, D_>), dim3(2 * rt().num_cus), blk, 0, \
                       st, a);                                                                   \
    break;
  switch (key) {
    PXR_FUSED_CASE(1, true, 0, 0) PXR_FUSED_CASE(1, true, 0, 3) PXR_FUSED_CASE(1, true, 1, 0)
    PXR_FUSED_CASE(1, true, 1, 3) PXR_FUSED_CASE(1, false, 0, 0) PXR_FUSED_CASE(1, false, 0, 3)
    PXR_FUSED_CASE(1, false, 1, 0) PXR_FUSED_CASE(1, false, 1, 3)
    PXR_FUSED_CASE(2, true, 0, 0) PXR_FUSED_CASE(2, true, 0, 3) PXR_FUSED_CASE(2, true, 1, 0)
    PXR_FUSED_CASE(2, true, 1, 3) PXR_FUSED_CASE(2, false, 0, 0) PXR_FUSED_CASE(2, false, 0, 3)
    PXR_FUSED_CASE(2, false, 1, 0) PXR_FUSED_CASE(2, false, 1, 3)
    PXR_FUSED_CASE(3, true, 0, 0) PXR_FUSED_CASE(3, true, 0, 3) PXR_FUSED_CASE(3, true, 1, 0)
    PXR_FUSED_CASE(3, true, 1, 3) PXR_FUSED_CASE(3, false, 0, 0) PXR_FUSED_CASE(3, false, 0, 3)
    PXR_FUSED_CASE(3, false, 1, 0) PXR_FUSED_CASE(3, false, 1, 3)
    default:
      return set_error(PXRHIP_ERR_ARG, "unsupported kernel combination");
  }
#undef PXR_FUSED_CASE
  PXR_LAUNCHED();
  return PXRHIP_OK;
}

}  // extern "C"
