/*
 * libnemo_setup.so -- host-side model set-up for NEMOLite2D in the
 * dl_esm_inf field layout. Plain C ABI, no GPU. It stands in for the
 * reference's Fortran model layer, which this library restates:
 *
 *   nemo_setup_tmask      common/model_mod.f90:94-249 (setup_tpoints_mask)
 *   nemo_setup_metrics    dl_esm_inf grid_init, as pinned by
 *                         original/nemolite2d.f90:218-231,274-290
 *   nemo_initialise       common/initialisation_mod.f90:7-63
 *   nemo_field_checksum   dl_esm_inf field_checksum (sum of |x| over the
 *                         internal region)
 *
 * Layout: every array is W x H, point (ji, jj) at jj*W + ji (0-based);
 * M x N is the T-mask extent (jpiglo x jpjglo, one boundary ring included),
 * the internal T region is ji = 1..M-2, jj = 1..N-2. W >= M+1 and H >= N+1
 * (the Fortran serial loops touch column M+1 and row N+1, 1-based); the
 * padding holds tmask 0.
 */
#ifndef NEMOLITE2D_SETUP_H
#define NEMOLITE2D_SETUP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary layouts. 1..3 are the reference's jphgr_msh values
 * (model_mod.f90:136-143); 0 and 4 have no reference counterpart. */
enum nemo_mask_mode {
    NEMO_MASK_CLOSED = 0,      /* all four edges solid */
    NEMO_MASK_SOUTH_OPEN = 1,  /* jphgr_msh = 1 */
    NEMO_MASK_EAST_OPEN = 2,   /* jphgr_msh = 2 */
    NEMO_MASK_SOUTH_EAST = 3,  /* jphgr_msh = 3 */
    NEMO_MASK_WEST_OPEN = 4    /* west column -1 for every row */
};

/* W = M+1 rounded up to a multiple of `align` (elements), H = N+1. */
void nemo_grid_dims(int M, int N, int align, int *W, int *H);

/* Fills tmask (W x H) for the single-domain case of setup_tpoints_mask:
 * 1 on the internal region, 0 elsewhere, then the open edges -1 (east before
 * south, so the south row wins at the corners, as in model_mod.f90:192-224).
 * Returns 0, or -1 for an unknown mode. */
int nemo_setup_tmask(int mode, int M, int N, int W, int H, int32_t *tmask);

/* Uniform Cartesian f-plane metrics over all W x H points: dx_* = dx,
 * dy_* = dy, area_* = dx*dy, gphiu = gphiv = gphi. */
void nemo_setup_metrics(int W, int H, double dx, double dy, double gphi,
                        double *area_t, double *area_u, double *area_v,
                        double *dx_u, double *dx_v, double *dx_t, double *dy_u,
                        double *dy_v, double *dy_t, double *gphiu,
                        double *gphiv);

/* Uniform depth (initialisation_mod.f90:26-28). */
void nemo_setup_depth_uniform(int W, int H, double depth, double *ht,
                              double *hu, double *hv);

/* Seeded smooth random bathymetry: SplitMix64 -> Box-Muller N(0,1) on the
 * internal T points (row-major), separable Gaussian low-pass (sigma cells,
 * truncated at 4 sigma, half-sample-symmetric edges), normalised to [-1,1]
 * by min/max over the internal region; ht = mean + relief*x clipped to
 * >= min_depth; points outside the internal region copy the nearest internal
 * value; hu = (ht(i)+ht(i+1))/2, hv = (ht(j)+ht(j+1))/2 (the last column /
 * row keep ht). Uses up to `threads` OpenMP threads (<= 0: all). */
void nemo_setup_bathymetry(uint64_t seed, int M, int N, int W, int H,
                           double sigma, double mean, double relief,
                           double min_depth, int threads, double *ht,
                           double *hu, double *hv);

/* initialisation_mod.f90:30-61: given sshn_t, derive sshn_u / sshn_v by the
 * area-weighted average over ji = 1..M-1 (jj = 1..N-1, 1-based) and zero
 * un, vn. Other arrays are left alone. */
void nemo_initialise(int M, int N, int W, int H, const double *area_t,
                     const double *area_u, const double *area_v,
                     const double *sshn_t, double *sshn_u, double *sshn_v,
                     double *un, double *vn);

/* Sum of |x| over the inclusive region, jj outer, ji inner. */
double nemo_field_checksum(const double *x, int W, int x0, int x1, int y0,
                           int y1);

/* SplitMix64 stream helper (the generator the bathymetry uses):
 * n uniforms in [0,1) from `seed`, 53-bit resolution. */
void nemo_uniforms(uint64_t seed, int64_t n, double *out);

/* ---- driver files (nemo_io.cpp) ---------------------------------------- */

/* One Fortran Ew.d field (w <= 32, d <= 17) into out[0..w), no terminator. */
void nemo_fortran_e(double x, int w, int d, char *out);

/* go2d_<istp>_<rank>.dat of common/gocean2d_io_mod.f90:98-149 (model_write)
 * to `path`: per internal T point (ji = xs..xe fastest, jj = ys..ye, 0-based)
 * '(6e16.7)' of x = ji*dx, y = jj*dy, ht, sshn, 0.5*(un(ji-1)+un(ji)),
 * 0.5*(vn(jj-1)+vn(jj)), and an empty record after each row. Returns 0, or
 * -1 when the file cannot be written. */
int nemo_write_results(const char *path, const double *ht, const double *sshn,
                       const double *un, const double *vn, int width, int xs, int xe,
                       int ys, int ye, double dx, double dy);

/* tmask_<rank>.dat of common/model_mod.f90:251-291 (dump_tmask): raw = 0,
 * '(2e16.7,1x,I3)' of x, y, tmask per internal T point; raw = 1,
 * '(3(I4,1x))' of the 1-based ji, jj and tmask over the whole M x N mask.
 * An empty record after each row. Returns 0 or -1. */
int nemo_dump_tmask(const char *path, const int *tmask, int width, int M, int N, int xs,
                    int xe, int ys, int ye, double dx, double dy, int raw);

#ifdef __cplusplus
}
#endif
#endif
