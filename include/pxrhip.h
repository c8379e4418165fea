/*
 * pxrhip.h -- C ABI of libpxrhip.so, the MI355X (gfx950) implementation of
 * PICSAR's per-step PIC hot path.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer (memory from hipMalloc or a
 *     torch ROCm tensor), except the small lo/hi bound triples and the FDTD
 *     stencil coefficient arrays, which are host pointers.
 *   - Particle arrays are contiguous float64 of length np.
 *   - Grid arrays are Fortran-ordered float64 with bounds
 *     (-ngx:nx+ngx, -ngy:ny+ngy, -ngz:nz+ngz): element (i,j,k) is at
 *     (i+ngx) + NX*((j+ngy) + NY*(k+ngz)), NX = nx+1+2*ngx, NY = ny+1+2*ngy.
 *   - Integer scalars are int64 (the reference's INTEGER(idp)); logicals are
 *     int64 0/1 (LOGICAL(lp)); floating scalars are double.
 *   - Arguments keep the reference routine's order.  Where the reference
 *     reads a module variable (fields%l_nodalgrid) it is passed explicitly as
 *     the trailing argument.
 *   - Outputs accumulate exactly where the reference accumulates (gathered
 *     fields, deposited currents).
 *   - Work is enqueued on the library's current stream (pxrhip_set_stream).
 *     With checks on (the default) a call that can find out-of-range
 *     particles synchronises that stream and reports them.
 *   - Every function returns PXRHIP_OK (0) or a negative error code;
 *     pxrhip_last_error() returns the message of the last failure.
 */
#ifndef PXRHIP_H
#define PXRHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PXRHIP_OK 0
#define PXRHIP_ERR_ARG (-1)         /* invalid argument (size, order, pointer) */
#define PXRHIP_ERR_HIP (-2)         /* a HIP runtime call failed */
#define PXRHIP_ERR_BOUNDS (-3)      /* particles outside the array (skipped) */
#define PXRHIP_ERR_UNSUPPORTED (-4) /* valid in the reference, not built here */
#define PXRHIP_ERR_CAPACITY (-5)    /* a buffer is too small for the request */

/* ---- runtime ------------------------------------------------------------ */
/* No counterpart in the reference (it has no device): library set-up. */
int pxrhip_init(int device);
int pxrhip_finalize(void);
const char *pxrhip_last_error(void);
int pxrhip_set_stream(void *hip_stream);
void *pxrhip_get_stream(void);
int pxrhip_synchronize(void);
/* on != 0: leaf calls synchronise and report out-of-range particles */
int pxrhip_set_checks(int on);
int pxrhip_version(void);
/* number of particles skipped as out of range since the last reset */
int pxrhip_bad_count(long long *count, int reset);

/* HIP graphs of the library's work (no counterpart in the reference).
 * pxrhip_init creates the library's own stream; pxrhip_graph_begin starts
 * capturing it, and every entry point called until pxrhip_graph_end is
 * recorded instead of run (only calls that do not synchronise may be
 * recorded: the tiled step, the field solvers and guard exchanges; not
 * copies to the host or checked leaf calls).  The recorded work keeps the
 * pointers it was given.  pxrhip_graph_launch replays it on the stream.
 * It also keeps the library's own scratch buffers and tile table: when a
 * later call grows a scratch buffer or switches to another grid
 * description, those are reallocated, pxrhip_graph_end fails if that
 * happened during the capture, and pxrhip_graph_launch of a graph captured
 * before fails with PXRHIP_ERR_ARG instead of replaying on freed memory. */
int pxrhip_graph_begin(void);
int pxrhip_graph_end(void **graph);
int pxrhip_graph_launch(void *graph);
int pxrhip_graph_destroy(void *graph);

/* ---- device memory, timers (no counterpart in the reference) ------------- */
/* The library runs on its own HIP runtime and stream.  Hosts that hold no
 * other GPU runtime allocate and move data here; copies to the host and
 * from pageable host memory synchronise the stream. */
int pxrhip_malloc(void **ptr, int64_t bytes);
int pxrhip_free(void *ptr);
int pxrhip_memcpy_htod(void *dst, const void *src, int64_t bytes);
int pxrhip_memcpy_dtoh(void *dst, const void *src, int64_t bytes);
int pxrhip_memcpy_dtod(void *dst, const void *src, int64_t bytes);
int pxrhip_fill_f64(double *p, int64_t n, double value);
int pxrhip_fill_i64(int64_t *p, int64_t n, int64_t value);
/* dst[i] = (double)src[i]: the weights of a mixed-precision species for
 * routines that take FP64 arrays (the charge density) */
int pxrhip_convert_f32_f64(const float *src, double *dst, int64_t n);
/* HIP events on the library stream (slots 0..65535), the device
 * counterpart of the reference's localtimes() wall-clock slots
 * (modules.F90:918-941) */
int pxrhip_timer_record(int slot);
int pxrhip_timer_elapsed(int start, int stop, double *ms);

/* ---- diagnostics --------------------------------------------------------- */
/* sum over np particles of (1/gaminv - 1)*w; times m*c^2 this is
 * compute_kin_energy (kin_energy.F90:33-143) for one species */
int pxrhip_kinetic_energy(int64_t np, const double *gaminv, const double *w,
                          double *out);
/* the same for a mixed-precision species (float32 momenta and w): the sum
 * of (gamma - 1)*w with gamma - 1 = u^2/(1 + sqrt(1 + u^2)) evaluated in
 * FP64 from the momenta, since a float32 gaminv cannot resolve 1/gaminv - 1
 * for a cold species (for u ~ 1e-3 it is a few float32 ulps). */
int pxrhip_kinetic_energy_f32(int64_t np, const float *ux, const float *uy,
                              const float *uz, const float *w, double *out);
/* sum of f^2 over the nodes 0..n-1 of each axis; times dV/2 this is one
 * component of the field energy of diags.F90:1072-1114 */
int pxrhip_field_energy(const double *f, int64_t nx, int64_t ny, int64_t nz,
                        int64_t ngx, int64_t ngy, int64_t ngz, double *out);
/* calc_field_div, diags.F90:129-156: divee = (ex(j)-ex(j-1))/dx +
 * (ey(k)-ey(k-1))/dy + (ez(l)-ez(l-1))/dz at the nodes 0..n of each axis;
 * the other nodes of divee are left as they are.  Guards >= 1. */
int pxrhip_calc_field_div(double *divee, const double *eex, const double *eey,
                          const double *eez, int64_t nx, int64_t ny, int64_t nz,
                          int64_t nxguard, int64_t nyguard, int64_t nzguard,
                          double dx, double dy, double dz);
/* calc_field_divB, diags.F90:170-197: forward differences of B at the
 * nodes 1..n of each axis.  Guards >= 1. */
int pxrhip_calc_field_divb(double *divbb, const double *bbx, const double *bby,
                           const double *bbz, int64_t nx, int64_t ny, int64_t nz,
                           int64_t nxguard, int64_t nyguard, int64_t nzguard,
                           double dx, double dy, double dz);
/* get_norm_divErho, diags.F90:1305-1345, on one domain: sqrt of the sum
 * over the nodes 1..n of each axis of (divee2*eps0 - rho2)^2, eps0 = 1
 * (normalised units). */
int pxrhip_get_norm_diverho(const double *divee2, const double *rho2, int64_t nx2,
                            int64_t ny2, int64_t nz2, int64_t nxguard,
                            int64_t nyguard, int64_t nzguard, double *norm);
/* get_loc_norm_2, diags.F90:1265-1293: the sum over the nodes 1..n of each
 * axis of array^2 (not square-rooted, as in the reference). */
int pxrhip_get_loc_norm_2(const double *array, int64_t nx2, int64_t ny2, int64_t nz2,
                          int64_t nxguard, int64_t nyguard, int64_t nzguard,
                          double *norm);

/* ---- field gathering ---------------------------------------------------- */
/* geteb3d_energy_conserving, field_gathering_manager_3d.F90:295-351.
 * Adds the Yee-staggered E and B at each particle into ex..bz.  Shape
 * orders 1..3 on each axis (others: PXRHIP_ERR_ARG, the reference has no
 * weights for them).  As geteb3d_energy_conserving_generic (:397-700)
 * dispatches: field_gathe_algo = 2, or orders other than 1_1_1, 2_2_2,
 * 3_3_3, run pxr_gete3d_n / pxr_getb3d_n_energy_conserving
 * (field_gathering_on_3d.F90:1042, 697), with their single-precision cubic
 * constants and ll4symtry; otherwise the order-N routines (lvect and the
 * other algorithms choose CPU code variants of the same sums).  l_nodal is
 * fields%l_nodalgrid (modules.F90). */
int pxrhip_geteb3d_energy_conserving(
    int64_t np, const double *xp, const double *yp, const double *zp,
    double *ex, double *ey, double *ez, double *bx, double *by, double *bz,
    double xmin, double ymin, double zmin, double dx, double dy, double dz,
    int64_t nx, int64_t ny, int64_t nz, int64_t nxguard, int64_t nyguard,
    int64_t nzguard, int64_t nox, int64_t noy, int64_t noz, const double *exg,
    const double *eyg, const double *ezg, const double *bxg, const double *byg,
    const double *bzg, int64_t ll4symtry, int64_t l_lower_order_in_v,
    int64_t lvect, int64_t field_gathe_algo, int64_t l_nodal);

/* ---- particle pushers --------------------------------------------------- */
/* pxr_boris_push_u_3d, boris_3d.F90:51-141 */
int pxrhip_boris_push_u_3d(int64_t np, double *uxp, double *uyp, double *uzp,
                           double *gaminv, const double *ex, const double *ey,
                           const double *ez, const double *bx, const double *by,
                           const double *bz, double q, double m, double dt);
/* pxr_ebcancelpush3d, vay_3d.F90:56-241 (which = 0 full, 1/2 half pushes) */
int pxrhip_ebcancelpush3d(int64_t np, double *uxp, double *uyp, double *uzp,
                          double *gi, const double *ex, const double *ey,
                          const double *ez, const double *bx, const double *by,
                          const double *bz, double q, double m, double dt,
                          int64_t which);
/* pxr_pushxyz, boris_3d.F90:849-893 */
int pxrhip_pushxyz(int64_t np, double *xp, double *yp, double *zp,
                   const double *uxp, const double *uyp, const double *uzp,
                   const double *gaminv, double dt);
/* pxr_set_gamma, boris_3d.F90:1059-1107 */
int pxrhip_set_gamma(int64_t np, const double *uxp, const double *uyp,
                     const double *uzp, double *gaminv);

/* ---- current deposition ------------------------------------------------- */
/* depose_jxjyjz, current_deposition_manager_3d.F90:47-67 ->
 * depose_jxjyjz_generic (:106-222).  current_depo_algo: 0 Esirkepov
 * (esirkepov_3d.F90:82/1734/3263), 1 the generic Esirkepov
 * pxr_depose_jxjyjz_esirkepov_n (:3480), 2 and 3 direct
 * (direct_current_deposition_3d.F90:99/879/1992).  Shape orders 0..3 on
 * each axis (others: PXRHIP_ERR_ARG); orders other than 1_1_1, 2_2_2, 3_3_3
 * run the generic Esirkepov whatever the algorithm, as in the reference.
 * Adds into jx, jy, jz.  Calls of many particles (tuning key
 * "leaf.bin_min") order them by 8^3-node block and sum each block in LDS
 * before adding it to the arrays; the result differs from the per-particle
 * form only in the order of the FP64 sums. */
int pxrhip_depose_jxjyjz(double *jx, double *jy, double *jz, int64_t np,
                         const double *xp, const double *yp, const double *zp,
                         const double *uxp, const double *uyp,
                         const double *uzp, const double *gaminv,
                         const double *w, double q, double xmin, double ymin,
                         double zmin, double dt, double dx, double dy,
                         double dz, int64_t nx, int64_t ny, int64_t nz,
                         int64_t nxguard, int64_t nyguard, int64_t nzguard,
                         int64_t nox, int64_t noy, int64_t noz,
                         int64_t current_depo_algo, int64_t l_nodal);

/* pxr_depose_rho_n, charge_deposition_3d.F90:2081-2228: charge density
 * with shape orders nox, noy, noz in 0..3 at the nodes (floor for odd, nint
 * for even orders), ADDED into rho.  l_particles_weight = 0 uses w[0] for
 * every particle; l4symtry folds x and y to their absolute values. */
int pxrhip_depose_rho_n(double *rho, int64_t np, const double *xp, const double *yp,
                        const double *zp, const double *w, double q, double xmin,
                        double ymin, double zmin, double dx, double dy, double dz,
                        int64_t nx, int64_t ny, int64_t nz, int64_t nxguard,
                        int64_t nyguard, int64_t nzguard, int64_t nox, int64_t noy,
                        int64_t noz, int64_t l_particles_weight, int64_t l4symtry);

/* ---- FDTD ---------------------------------------------------------------- */
/* pxrpush_em3d_evec, yee.F90:637-750.  lo/hi triples are host int64[3];
 * each array carries its own bounds, as in the reference. */
int pxrhip_push_em3d_evec(
    const int64_t *xlo, const int64_t *xhi, const int64_t *ylo,
    const int64_t *yhi, const int64_t *zlo, const int64_t *zhi, double *ex,
    const int64_t *exlo, const int64_t *exhi, double *ey, const int64_t *eylo,
    const int64_t *eyhi, double *ez, const int64_t *ezlo, const int64_t *ezhi,
    const double *bx, const int64_t *bxlo, const int64_t *bxhi,
    const double *by, const int64_t *bylo, const int64_t *byhi,
    const double *bz, const int64_t *bzlo, const int64_t *bzhi,
    const double *jx, const int64_t *jxlo, const int64_t *jxhi,
    const double *jy, const int64_t *jylo, const int64_t *jyhi,
    const double *jz, const int64_t *jzlo, const int64_t *jzhi, double mudt,
    double dtsdx, double dtsdy, double dtsdz);
/* pxrpush_em3d_bvec, yee.F90:1258-1370 */
int pxrhip_push_em3d_bvec(
    const int64_t *xlo, const int64_t *xhi, const int64_t *ylo,
    const int64_t *yhi, const int64_t *zlo, const int64_t *zhi,
    const double *ex, const int64_t *exlo, const int64_t *exhi,
    const double *ey, const int64_t *eylo, const int64_t *eyhi,
    const double *ez, const int64_t *ezlo, const int64_t *ezhi, double *bx,
    const int64_t *bxlo, const int64_t *bxhi, double *by, const int64_t *bylo,
    const int64_t *byhi, double *bz, const int64_t *bzlo, const int64_t *bzhi,
    double dtsdx, double dtsdy, double dtsdz);
/* pxrpush_em3d_evec_norder, yee.F90:46-133; dtsdx..dtsdz are host arrays of
 * norder/2 coefficients (norder even, <= 32). */
int pxrhip_push_em3d_evec_norder(
    double *ex, double *ey, double *ez, const double *bx, const double *by,
    const double *bz, const double *jx, const double *jy, const double *jz,
    double mudt, const double *dtsdx, const double *dtsdy,
    const double *dtsdz, int64_t nx, int64_t ny, int64_t nz, int64_t norderx,
    int64_t nordery, int64_t norderz, int64_t nxguard, int64_t nyguard,
    int64_t nzguard, int64_t nxs, int64_t nys, int64_t nzs,
    int64_t l_nodalgrid);
/* pxrpush_em3d_bvec_norder, yee.F90:752-833 */
int pxrhip_push_em3d_bvec_norder(
    const double *ex, const double *ey, const double *ez, double *bx,
    double *by, double *bz, const double *dtsdx, const double *dtsdy,
    const double *dtsdz, int64_t nx, int64_t ny, int64_t nz, int64_t norderx,
    int64_t nordery, int64_t norderz, int64_t nxguard, int64_t nyguard,
    int64_t nzguard, int64_t nxs, int64_t nys, int64_t nzs,
    int64_t l_nodalgrid);

/* ---- guard handling on one periodic rank -------------------------------- */
/* field_bc -> exchange_mpi_3d_grid_array_with_guards
 * (field_boundaries.F90:81-260) when every neighbour is the rank itself. */
int pxrhip_field_bc_periodic(double *field, int64_t nx, int64_t ny, int64_t nz,
                             int64_t nxg, int64_t nyg, int64_t nzg);
/* summation_bcs (field_boundaries.F90:403-500) on one periodic rank. */
int pxrhip_current_bc_periodic(double *array, int64_t nx, int64_t ny,
                               int64_t nz, int64_t nxg, int64_t nyg,
                               int64_t nzg);

/* ---- sorting ------------------------------------------------------------- */
/* pxr_particle_bin_sorting, sorting.F90:247-361: stable counting sort of
 * np2 particles by sorting-grid cell; pid points at the first pid column
 * (np2 values), the only one the reference permutes.  Keys outside
 * [1, nx3*ny3*nz3] are reported (PXRHIP_ERR_BOUNDS) and nothing is moved. */
int pxrhip_particle_bin_sorting(int64_t np2, double *xp, double *yp,
                                double *zp, double *ux, double *uy, double *uz,
                                double *gam, double *pid, int64_t wpid,
                                double xmin2, double ymin2, double zmin2,
                                double xmax2, double ymax2, double zmax2,
                                double dxf, double dyf, double dzf);

/* No counterpart in the reference.  pxrhip_particle_bin_sorting for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * loops, harnesses that forbid hipMalloc): the same arguments, the same
 * result, but its keys, histogram, permutation and staging live in a fixed
 * device array of the library (about 20 MB per GPU, held until the process
 * ends) instead of buffers it allocates.  Limits: np2 <= 2^20 particles and
 * at most 2^20 sorting-grid bins, else PXRHIP_ERR_ARG; pxrhip_init must have
 * been called (it allocates, so this entry point does not call it).  Every
 * call on one GPU uses that GPU's one array: calls that overlap on a GPU
 * (other host threads, other streams) race. */
int pxrhip_particle_bin_sorting_noalloc(int64_t np2, double *xp, double *yp,
                                        double *zp, double *ux, double *uy,
                                        double *uz, double *gam, double *pid,
                                        int64_t wpid, double xmin2,
                                        double ymin2, double zmin2,
                                        double xmax2, double ymax2,
                                        double zmax2, double dxf, double dyf,
                                        double dzf);

/* ---- tiled step (the managers' work on device-resident tiles) ----------- */
/* Geometry shared by the tile-level entry points: one periodic rank with
 * nx*ny*nz cells, field and current arrays of identical shape with guards
 * ngx..ngz, split into ntx*nty*ntz tiles as set_tile_split_for_species
 * does (tiling.F90:94-271): n0 = nx_grid/ntx nodes per tile, the last tile
 * takes the remainder.  Particles of one species are stored SoA, packed by
 * tile: tile t holds [tile_off[t], tile_off[t] + tile_np[t]). */
typedef struct {
  int64_t nx, ny, nz;    /* cells */
  int64_t ngx, ngy, ngz; /* guard nodes of every grid array */
  int64_t ntx, nty, ntz; /* tiles (each >= 4 nodes wide, as the reference requires) */
  double xmin, ymin, zmin, dx, dy, dz;
  /* per-tile sorting grid: cell sizes (absolute, i.e. the input value times
   * dx as submain.F90:601 does) and the shift of its origin from the
   * tile's lower edge (particle_sorting_sub, sorting.F90:150-160) */
  double sorting_dx, sorting_dy, sorting_dz;
  double sorting_shiftx, sorting_shifty, sorting_shiftz;
  /* particle boundary at x_min, x_max, y_min, y_max, z_min, z_max
   * (pbound_x_min ..., modules.F90): PXRHIP_PBOUND_* */
  int64_t pbound[6];
  /* the particle exchange of particle_bcs (particle_boundaries.F90:83-217,
   * the reference's partcom): 0, its default (control_file.F90:139), is
   * particle_bcs_tiles_and_mpi_3d (:2107-3031), whose domain boundaries
   * know absorbing and reinjecting faces and treat every other value,
   * reflecting included, as periodic (:2308-2333); 1 and 2 are
   * particle_bcs_tiles followed by particle_bcs_mpi_blocking (:1040-1371,
   * mpicom_curr = 1, its default), where a reflecting face moves the
   * particle one cell back in and reverses its normal momentum. */
  int64_t partcom;
} pxrhip_grid;

#define PXRHIP_PBOUND_PERIODIC 0
#define PXRHIP_PBOUND_ABSORBING 1   /* the particle is removed */
#define PXRHIP_PBOUND_REFLECTING 2  /* partcom 1, 2: one cell back in, normal momentum
                                       reversed; partcom 0: periodic */
#define PXRHIP_PBOUND_REINJECTING 3 /* mirrored in, at rest, kept in its tile */

#define PXRHIP_MAX_PID_EXTRA 4

typedef struct {
  double *x, *y, *z, *ux, *uy, *uz, *gaminv, *w; /* SoA particle arrays */
  int64_t *tile_off;  /* device, ntiles+1 entries: packed tile offsets, last = total */
  int64_t *tile_np;   /* device, ntiles entries */
  int64_t *bin_start; /* device, nbins+1 entries: first particle of every
                         sorting-grid cell of every tile (filled by
                         pxrhip_tiles_rebin_sort), or NULL */
  int64_t capacity;   /* allocated length of each particle array */
  double q, m;
  /* 0: every particle array is float64.  1: mixed precision -- x, y, z are
   * float64, while ux, uy, uz, gaminv and w point to float32 arrays (the
   * pointers are cast); gather, push and the LDS current are FP32 and the
   * global current is FP64. */
  int64_t precision;
  /* Columns 2 .. npid_extra+1 of the reference's pid array (particle_tile%pid,
   * modules.F90:339-452 with npid, ssnpid, xoldpid at :608-617: particle
   * ids, old positions ...), float64 whatever the precision, capacity
   * values each.  The re-binning carries them with their particles (as
   * particle_bcs_tiles does); no kernel reads them.  Column 1, the weight,
   * is w.  npid_extra is 0 .. PXRHIP_MAX_PID_EXTRA. */
  int64_t npid_extra;
  double *pid_extra[PXRHIP_MAX_PID_EXTRA];
} pxrhip_species;

/* Laser antenna particles: laserp_pusher_gaussian
 * (laser_pusher_manager_3d.F90:195-271, profile gaussian_profile :368-415)
 * and laserp_pusher_hanning (:281-356, hanning_profile :426-478), same
 * argument order.  pid is the device (np x npidd) column-major pid array;
 * columns 2 and 3 hold the antenna-frame coordinates.  polvector1,
 * polvector2 (3 doubles each) and q_z (complex: re, im) are host arrays.
 * u = charge * amp * 0.01 c / emax, gaminv = 1, x += dtt * u (the
 * reference reads the module's dt, which its caller passes as dtt).
 * lvect (a CPU blocking factor) is ignored.  Normalised units, c = 1. */
int pxrhip_laserp_pusher_gaussian(
    int64_t np, int64_t npidd, const double *pid, double *xp, double *yp,
    double *zp, double *uxp, double *uyp, double *uzp, double *gaminv,
    double dtt, int64_t lvect, double emax, double emax1, double emax2,
    const double *polvector1, const double *polvector2, double k0_laser,
    const double *q_z, double charge, double laser_tau, double real_time,
    double t_peak, int64_t temporal_order, double polangle);
int pxrhip_laserp_pusher_hanning(
    int64_t np, int64_t npidd, const double *pid, double *xp, double *yp,
    double *zp, double *uxp, double *uyp, double *uzp, double *gaminv,
    double dtt, int64_t lvect, double emax, double emax1, double emax2,
    const double *polvector1, const double *polvector2, double k0_laser,
    const double *q_z, double charge, double real_time, double t_peak,
    int64_t temporal_order, double polangle);

/* Number of sorting-grid cells over all tiles (the length of bin_start
 * minus one) for this geometry. */
int pxrhip_tiles_nbins(const pxrhip_grid *grid, int64_t *nbins);

/* One fused pass over every particle of one species:
 * energy-conserving gather (field_gathering_plus_particle_pusher_*,
 * particle_pusher_manager_3d.F90:431-672) -> pusher (0 Boris + set_gamma,
 * 1 Vay) -> pxr_pushxyz -> current deposition of the new positions
 * (current_deposition_manager_3d.F90:224-505), accumulated into jx..jz.
 * The stencil blocks of each chunk of particles are staged in LDS.
 * depo_algo: 0 Esirkepov, 3 direct.  chunk: particles per work item
 * (0 = library default). */
int pxrhip_tiles_gather_push_deposit(
    const pxrhip_grid *grid, const pxrhip_species *sp, const double *exg,
    const double *eyg, const double *ezg, const double *bxg, const double *byg,
    const double *bzg, double *jx, double *jy, double *jz, double dt,
    int64_t order, int64_t l_lower_order_in_v, int64_t pusher,
    int64_t depo_algo, int64_t chunk);

/* field_gathering_plus_particle_pusher (particle_pusher_manager_3d.F90:
 * 431-672) on a tiled species: the gather, push and position push of
 * pxrhip_tiles_gather_push_deposit without its deposition (pusher 0 Boris,
 * 1 Vay).  The species should be binned; particles that moved since still
 * gather, from global memory. */
int pxrhip_tiles_gather_push(
    const pxrhip_grid *grid, const pxrhip_species *sp, const double *exg,
    const double *eyg, const double *ezg, const double *bxg, const double *byg,
    const double *bzg, double dt, int64_t order, int64_t l_lower_order_in_v,
    int64_t pusher, int64_t chunk);

/* The current of a binned species, ADDED into jx..jz: the per-species work
 * of pxrdepose_currents_on_grid_jxjyjz (current_deposition_manager_3d.F90:
 * 224-1830) with depose_jxjyjz's shapes (depo_algo 0 Esirkepov, 3 direct,
 * order 1..3), from the positions and momenta as they stand -- after
 * particle_bcs, as submain.F90:125-140 orders the step.  The current
 * blocks of the work items are accumulated in LDS. */
int pxrhip_tiles_depose_currents(
    const pxrhip_grid *grid, const pxrhip_species *sp, double *jx, double *jy,
    double *jz, double dt, int64_t order, int64_t depo_algo, int64_t chunk);

/* particle_bcs on one rank (particle_boundaries.F90:83-426 and the domain
 * boundaries of the exchange that grid->partcom selects, per grid->pbound;
 * absorbed particles are dropped, so dst may hold fewer) followed by
 * particle_sorting_sub (sorting.F90:106-215): every particle that left the
 * box is wrapped, reflected, reinjected or removed (in src, in place), then
 * assigned to the tile that now holds it (tile index of
 * particle_boundaries.F90:397) and the tiles are ordered by sorting-grid
 * cell (pxr_particle_bin_sorting's key).  The result is written to dst,
 * which must have the same capacity; dst's tile_off, tile_np and bin_start
 * are filled in (tile_off[ntiles] is the number of particles kept).
 * The sort is stable: the particles of one cell keep their order in src,
 * as pxr_particle_bin_sorting's counting sort keeps it (sorting.F90:
 * 333-343), so the result does not depend on the order in which GPU waves
 * ran (tuning key "rebin.stable" 0 drops that order for a cheaper scatter). */
int pxrhip_tiles_rebin_sort(const pxrhip_grid *grid, pxrhip_species *src,
                            pxrhip_species *dst);

/* Charge density of a tiled species, ADDED into rho (guards included, not
 * folded): the per-tile blocks of pxrdepose_rho_on_grid_sub_openmp_3d
 * (charge_deposition_manager.F90:516-743) with the shapes of
 * pxr_depose_rho_n (charge_deposition_3d.F90:2081-2228) of order 1..3.  The
 * species must be binned (as pxrhip_tiles_rebin_sort leaves it); particles
 * that moved since are still deposited, from global memory.  Particles
 * outside the arrays are skipped and counted for pxrhip_tiles_check.  A
 * mixed-precision species' float32 weights are read as such. */
int pxrhip_tiles_depose_rho(const pxrhip_grid *grid, const pxrhip_species *sp,
                            double *rho, int64_t order);

/* Launch parameters of the tiled kernels.  No counterpart in the reference
 * (its CPU blocking factors lvec_* play this role).  Keys:
 * "fused.lds_bytes" (dynamic LDS per workgroup), "fused.wg_per_cu",
 * "fused.spread" (0/1, lane-spread particle order in the deposition phase),
 * "fused.sstride" (largest lane stride of that order, 0: a sixty-fourth of
 * the work item; -1, the default, sets one per work item from its particles
 * per sorting cell and the shape order, none at order 3) and
 * "mixed.sstride" (the same for a mixed-precision species, whose order-3
 * work items are capped too),
 * "fused.chunk" (default particles per work item; a work item also holds no
 * more sorting-grid rows than its LDS blocks allow),
 * "fused.j_lds_bytes" (largest current block accumulated in LDS; negative:
 * all of fused.lds_bytes), "fused.jpad" (0/1, odd row and plane pitches of
 * the LDS current block), "fused.margin_pct" (motion margin, in 1/100 cell,
 * of the current block estimated when work items are cut), "mixed.jmargin"
 * (nodes of motion room per side that the work items of a mixed-precision
 * species with the fixed-point current leave around the current block of
 * their particles' old positions: -1, the default, takes 0 with the
 * Esirkepov deposition and 1 with the direct one; a work item whose pushed
 * particles outgrow its LDS block is deposited by a second kernel through a
 * block cut down to fit, the particles outside it adding into global
 * memory; other species keep 1), "mixed.jacc"
 * (LDS current of a mixed-precision species: 0 FP32 atomics, 1 FP32
 * atomics compensated by TwoSum into a second array, 2 FP64 atomics, 3 the
 * default: 64-bit fixed point), "fused.jacc" (LDS current of an FP64
 * species: 0 FP64 atomics, 3 the default: 64-bit fixed point),
 * "fused.order" (0, the default: work items in the order they are built;
 * 1 largest first), "mixed.lds_bytes" and "mixed.wg_per_cu" (the launch of
 * a mixed-precision species; fused.lds_bytes and fused.wg_per_cu are those
 * of an FP64 one), "rebin.stable" (1, the default: pxrhip_tiles_rebin_sort
 * keeps the source order inside a cell; 0: the order in which the GPU's
 * waves arrive), "leaf.bin_min" (pxrhip_depose_jxjyjz with the Esirkepov
 * current and pxrhip_depose_rho_n sum calls of at least this many particles,
 * and one per 8^3-node block of the arrays, block by block in LDS; default
 * 65536, 0: never) and "leaf.bin_chunk" (largest work item of that form,
 * default 4096 particles).
 * wg_per_cu 0 (the default) is automatic: three workgroups per CU with
 * 52224 bytes each when a work item of four sorting rows of the widest tile
 * fits that, else two with 79616 bytes; a wg_per_cu above 0 takes the
 * matching lds_bytes (default 52224).  An FP64 species at order 3 with the
 * direct deposition runs at most two workgroups per CU, since its kernel
 * holds registers for two waves per SIMD; it then takes 2 and 79616 bytes.
 * pxrhip_get_tuning also reads "last.wg_per_cu" and "last.lds_bytes", the
 * launch of the last pxrhip_tiles_gather_push_deposit, and
 * "device.num_cus".  Work items whose
 * field or current block exceeds its LDS budget are completed from global
 * memory by a second kernel (a mixed-precision species' current: through
 * an LDS block cut down to fit, see "mixed.jmargin"). */
int pxrhip_set_tuning(const char *key, int64_t value);
int pxrhip_get_tuning(const char *key, int64_t *value);

/* Work items of the last pxrhip_tiles_gather_push_deposit and how many of
 * them did not fit in LDS: run from global memory, or deposited through a
 * cut-down current block (synchronises). */
int pxrhip_tiles_stats(int64_t *nitems, int64_t *noverflow);

/* Reads the device-side error counter of the tiled path (particles found
 * outside the arrays and skipped) and resets it; a non-zero count is
 * returned as PXRHIP_ERR_BOUNDS. */
int pxrhip_tiles_check(void);

#ifdef __cplusplus
}
#endif
#endif /* PXRHIP_H */
