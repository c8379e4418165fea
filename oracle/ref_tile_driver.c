/* ref_tile_driver.c -- TEST INFRASTRUCTURE: the CPU baseline's tile loop.
 *
 * Runs the reference's own compiled FP64 kernels (libpxrref_o3.so, built by
 * build_ref.sh from the PICSAR sources with its production flags) over
 * tiles with OpenMP threads, as the reference's managers do: one tile per
 * iteration of an OpenMP loop, and inside it, for each species,
 *   - the order-3 energy-conserving gather (geteb3d_energy_conserving_vecv3
 *     _3_3_3, field_gathering_o3_3d.F90:5611; what the default algo of
 *     field_gathering_plus_particle_pusher runs),
 *   - the Vay push (pxr_ebcancelpush3d, vay_3d.F90:56) and pxr_pushxyz
 *     (boris_3d.F90:849),
 *   - the order-3 Esirkepov deposition into the tile's J block
 *     (depose_jxjyjz_esirkepov_3_3_3, esirkepov_3d.F90:3263), the block
 *     zeroed first as current_deposition_manager_3d.F90:589 does.
 * No Python runs inside the timed loop.  bench.py's cpu_baseline leg is the
 * only caller; it never measures or ships the product through here.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t I;

/* Fortran entry points of the reference (every argument by reference) */
void geteb3d_energy_conserving_vecv3_3_3_3_(
    I *np, double *xp, double *yp, double *zp, double *ex, double *ey, double *ez, double *bx,
    double *by, double *bz, double *xmin, double *ymin, double *zmin, double *dx, double *dy,
    double *dz, const double *exg, I *exg_ng, I *exg_nv, const double *eyg, I *eyg_ng,
    I *eyg_nv, const double *ezg, I *ezg_ng, I *ezg_nv, const double *bxg, I *bxg_ng,
    I *bxg_nv, const double *byg, I *byg_ng, I *byg_nv, const double *bzg, I *bzg_ng,
    I *bzg_nv, I *lvect, I *l_lower_order_in_v, I *l_nodal);
void pxr_ebcancelpush3d_(I *np, double *ux, double *uy, double *uz, double *gi, double *ex,
                         double *ey, double *ez, double *bx, double *by, double *bz, double *q,
                         double *m, double *dt, I *which);
void pxr_pushxyz_(I *np, double *x, double *y, double *z, double *ux, double *uy, double *uz,
                  double *gi, double *dt);
void depose_jxjyjz_esirkepov_3_3_3_(double *jx, I *jx_ng, I *jx_nv, double *jy, I *jy_ng,
                                    I *jy_nv, double *jz, I *jz_ng, I *jz_nv, I *np, double *xp,
                                    double *yp, double *zp, double *ux, double *uy, double *uz,
                                    double *gi, double *w, double *q, double *xmin, double *ymin,
                                    double *zmin, double *dt, double *dx, double *dy, double *dz);

/* Time `seconds` of tile work on `nthreads` threads.  Tile geometry: nc
 * cells per axis, ng guards, fields[6] the E and B blocks of one tile
 * ((nc+1+2ng)^3 each, Fortran order).  pool holds, for each thread, two
 * species x 8 arrays (x, y, z, ux, uy, uz, gaminv, w) of ntile particles,
 * copied into the thread's work arrays before each tile, as the tile's
 * particles would be read from memory.  q[2], m[2]: the species.
 * Returns the particle-steps done in *steps and the wall time in *wall. */
int pxrref_time_tiles(I nthreads, double seconds, I nc, I ng, I ntile, const double *const *fields,
                      const double *pool, const double *q, const double *m, double dt,
                      double *steps, double *wall) {
  if (nthreads < 1 || nc < 1 || ng < 0 || ntile < 1 || !fields || !pool || !steps || !wall)
    return -1;
  const I nn = nc + 1 + 2 * ng, nblk = nn * nn * nn;
  double total = 0.0;
  const double t0 = omp_get_wtime();
  double tend = t0;
#pragma omp parallel num_threads((int)nthreads) reduction(+ : total) reduction(max : tend)
  {
    const int t = omp_get_thread_num();
    double *work = (double *)malloc(sizeof(double) * (size_t)(ntile * 14 + 3 * nblk));
    if (work) {
      double *p[8], *e[6], *J[3];
      for (int k = 0; k < 8; k++) p[k] = work + (size_t)k * ntile;
      for (int k = 0; k < 6; k++) e[k] = work + (size_t)(8 + k) * ntile;
      for (int k = 0; k < 3; k++) J[k] = work + (size_t)14 * ntile + (size_t)k * nblk;
      I ngv[3] = {ng, ng, ng}, nv[3] = {nc + 1, nc + 1, nc + 1};
      I np = ntile, lvect = 64, lo = 1, nod = 0, which = 0;
      double zero = 0.0, one = 1.0, dtl = dt;
      while (omp_get_wtime() - t0 < seconds) {
        for (int s = 0; s < 2; s++) {
          const double *src = pool + ((size_t)t * 2 + s) * 8 * (size_t)ntile;
          memcpy(work, src, sizeof(double) * 8 * (size_t)ntile);
          memset(e[0], 0, sizeof(double) * 6 * (size_t)ntile);
          memset(J[0], 0, sizeof(double) * 3 * (size_t)nblk);
          double qs = q[s], ms = m[s];
          geteb3d_energy_conserving_vecv3_3_3_3_(
              &np, p[0], p[1], p[2], e[0], e[1], e[2], e[3], e[4], e[5], &zero, &zero, &zero,
              &one, &one, &one, fields[0], ngv, nv, fields[1], ngv, nv, fields[2], ngv, nv,
              fields[3], ngv, nv, fields[4], ngv, nv, fields[5], ngv, nv, &lvect, &lo, &nod);
          pxr_ebcancelpush3d_(&np, p[3], p[4], p[5], p[6], e[0], e[1], e[2], e[3], e[4], e[5],
                              &qs, &ms, &dtl, &which);
          pxr_pushxyz_(&np, p[0], p[1], p[2], p[3], p[4], p[5], p[6], &dtl);
          depose_jxjyjz_esirkepov_3_3_3_(J[0], ngv, nv, J[1], ngv, nv, J[2], ngv, nv, &np, p[0],
                                         p[1], p[2], p[3], p[4], p[5], p[6], p[7], &qs, &zero,
                                         &zero, &zero, &dtl, &one, &one, &one);
          total += (double)ntile;
        }
      }
      free(work);
    }
    tend = omp_get_wtime();
  }
  *steps = total;
  *wall = tend - t0;
  return 0;
}
