/*
 * gray_spacetime.h -- numerical-spacetime inputs (part of libgray.so):
 * a boosted Kerr-Schild sampler (the role of the reference's
 * tools/generate_data.py + tools/boosted_ks.py) and HDF5 files in the
 * reference's layout (README.md:40-60, sim-org/io.c:149-533):
 *   /grid/x, /grid/y, /grid/z        1-D axes
 *   /<time>/Gamma_<a><b><c>          40 datasets [nz][ny][nx], a in txyz,
 *                                     b <= c in txyz
 *   /<time>/g_<a><b>                 10 datasets, a <= b
 *   /<time>/rho                      1 dataset
 * Every group not named with a leading 'g' is a time level; its name,
 * read with strtod, is its time.
 */
#ifndef GRAY_SPACETIME_H
#define GRAY_SPACETIME_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boosted Kerr-Schild black hole: mass m, spin a, moving with speed v
 * along the unit vector n. */
typedef struct {
	double m, a, v;
	double n[3];
} gray_bks;

void gray_bks_setup(gray_bks *b, double m, double a, double v, const double axis[3]);
/* Metric met[10] (tt tx ty tz xx xy xz yy yz zz) and, if gam is not NULL,
 * Christoffels gam[40] (Gamma^mu_{ab}, mu-major, ab in the same order) at
 * the event x[4] = (t, x, y, z). */
void gray_bks_point(const gray_bks *b, const double x[4], double met[10], double gam[40]);
/* Sample on the grid xs[nx] x ys[ny] x zs[nz] at time t into interleaved
 * per-point arrays (p = (k*ny + j)*nx + i): gam[p*40 + c], met[p*10 + c],
 * rho[p] = 1/|x| (any may be NULL).  OpenMP over points. */
void gray_bks_sample(const gray_bks *b, double t, int nx, int ny, int nz, const double *xs,
                     const double *ys, const double *zs, double *gam, double *met, double *rho);
void gray_axis_uniform(double lo, double hi, int n, double *out);
void gray_axis_fisheye(double lo, double hi, int n, double *out);

/* ---- HDF5 files (all return 0 on success, -1 on failure with a message
 * in gray_h5_error()) ---- */
const char *gray_h5_error(void);
/* Create a file with the three axes (stored in `dtype`, 4 or 8 bytes). */
int gray_h5_create(const char *path, int nx, int ny, int nz, const double *xs,
                   const double *ys, const double *zs, int dtype);
/* Append one time level named `name` from interleaved FP64 arrays. */
int gray_h5_write_level(const char *path, const char *name, int nx, int ny, int nz,
                        const double *gam, const double *met, const double *rho, int dtype);
/* The rewrite's ray dump (replaces sim/gray.c:170-176): create `path`
 * (error if it exists) with dataset /rays of dims {W, H, nd} from states in
 * dump order [H][W][nd], stored as dtype (4 or 8) bytes per real. */
int gray_h5_write_rays(const char *path, int w, int h, int nd, const double *data, int dtype);
/* Read /rays: dims into dims[3]; values (if out is not NULL) as FP64 in file
 * order [W][H][nd]. */
int gray_h5_read_rays(const char *path, int dims[3], double *out);
/* Time levels: names (sorted by time) into names[max][64], times[max];
 * returns the count or -1 (fewer than 2 groups is an error, io.c:182-185). */
int gray_h5_list_times(const char *path, char (*names)[64], double *times, int max);
/* Axes: n[3] always; xs/ys/zs (may be NULL) receive the values; *dtype the
 * stored element size. */
int gray_h5_read_axes(const char *path, int n[3], double *xs, double *ys, double *zs,
                      int *dtype);
/* One level into interleaved arrays of `out_dtype` reals (4 or 8):
 * gam[p*40 + c], met[p*10 + c]; rho (may be NULL) n_points reals.  Every
 * dataset must hold nx*ny*nz values (io.c:377-380). */
int gray_h5_read_level(const char *path, const char *name, const int n[3], void *gam,
                       void *met, void *rho, int out_dtype);
/* Map of an axis: 0 uniform, 1 fisheye (sinh((A i + B)^3)), -1 neither. */
int gray_axis_map(const double *x, int n);

#ifdef __cplusplus
}
#endif
#endif
