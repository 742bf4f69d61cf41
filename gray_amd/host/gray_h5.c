/*
 * gray_h5.c -- HDF5 files of numerical spacetimes in the reference's layout
 * (README.md:40-60; reader sim-org/io.c:149-533; writer
 * tools/generate_data.py:132-160).  Built into libgray_h5.so against the
 * HDF5 C library (1.10).
 *
 * Differences from the reference reader, each a decision for behaviour it
 * leaves undefined:
 *   - datasets are read in their native precision and converted, so FP32
 *     and FP64 files both load (io.c:434-435 always makes FP32 images);
 *   - axis values are read in full and the coordinate map (uniform or the
 *     fisheye sinh((A i + B)^3)) is detected from them (interp.cl
 *     hard-codes fisheye from the end points only).
 */
#define _GNU_SOURCE
#include "../../include/gray_spacetime.h"

#include <hdf5.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char h5err[512];

static int h5fail(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(h5err, sizeof(h5err), fmt, ap);
	va_end(ap);
	return -1;
}

const char *gray_h5_error(void) { return h5err; }

static const char dim[4] = {'t', 'x', 'y', 'z'};

/* Dataset names in the reader's order: Gamma_ijk (i, j <= k), then g_ij
 * (i <= j), io.c:356-361 and :390-393. */
static void gamma_name(int c, char *buf, size_t n)
{
	int mu = c / 10, r = c % 10, a = 0, b = 0, k = 0;
	for (a = 0; a < 4; ++a)
		for (b = a; b < 4; ++b, ++k)
			if (k == r) {
				snprintf(buf, n, "Gamma_%c%c%c", dim[mu], dim[a], dim[b]);
				return;
			}
}

static void metric_name(int c, char *buf, size_t n)
{
	int a, b, k = 0;
	for (a = 0; a < 4; ++a)
		for (b = a; b < 4; ++b, ++k)
			if (k == c) {
				snprintf(buf, n, "g_%c%c", dim[a], dim[b]);
				return;
			}
}

static hid_t file_type(int dtype) { return dtype == 4 ? H5T_IEEE_F32LE : H5T_IEEE_F64LE; }

static int write_1d(hid_t loc, const char *name, const double *v, hsize_t n, int dtype)
{
	hid_t sp = H5Screate_simple(1, &n, NULL);
	hid_t ds = H5Dcreate2(loc, name, file_type(dtype), sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
	herr_t st = ds < 0 ? -1 : H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
	if (ds >= 0)
		H5Dclose(ds);
	H5Sclose(sp);
	return st < 0 ? -1 : 0;
}

int gray_h5_create(const char *path, int nx, int ny, int nz, const double *xs,
                   const double *ys, const double *zs, int dtype)
{
	hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
	if (f < 0)
		return h5fail("cannot create %s", path);
	hid_t g = H5Gcreate2(f, "grid", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
	int rc = g < 0 || write_1d(g, "x", xs, (hsize_t)nx, dtype) ||
	         write_1d(g, "y", ys, (hsize_t)ny, dtype) || write_1d(g, "z", zs, (hsize_t)nz, dtype);
	if (g >= 0)
		H5Gclose(g);
	H5Fclose(f);
	return rc ? h5fail("writing the axes of %s failed", path) : 0;
}

int gray_h5_write_level(const char *path, const char *name, int nx, int ny, int nz,
                        const double *gam, const double *met, const double *rho, int dtype)
{
	const size_t np = (size_t)nx * ny * nz;
	hsize_t dims[3] = {(hsize_t)nz, (hsize_t)ny, (hsize_t)nx};
	double *buf = malloc(np * sizeof(double));
	if (!buf)
		return h5fail("out of memory");
	hid_t f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
	if (f < 0) {
		free(buf);
		return h5fail("cannot open %s", path);
	}
	hid_t g = H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
	hid_t sp = H5Screate_simple(3, dims, NULL);
	int rc = g < 0 || sp < 0;
	for (int c = 0; !rc && c < 51; ++c) {
		char dn[32];
		const double *src;
		int stride, off;
		if (c < 40) {
			gamma_name(c, dn, sizeof(dn));
			src = gam; stride = 40; off = c;
		} else if (c < 50) {
			metric_name(c - 40, dn, sizeof(dn));
			src = met; stride = 10; off = c - 40;
		} else {
			snprintf(dn, sizeof(dn), "rho");
			src = rho; stride = 1; off = 0;
		}
		if (!src)
			continue;
		for (size_t p = 0; p < np; ++p)
			buf[p] = src[p * stride + off];
		hid_t ds = H5Dcreate2(g, dn, file_type(dtype), sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
		rc = ds < 0 || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0;
		if (ds >= 0)
			H5Dclose(ds);
	}
	if (sp >= 0)
		H5Sclose(sp);
	if (g >= 0)
		H5Gclose(g);
	H5Fclose(f);
	free(buf);
	return rc ? h5fail("writing level %s of %s failed", name, path) : 0;
}

typedef struct {
	char (*names)[64];
	int n, max;
} name_list;

static herr_t collect(hid_t g, const char *name, const H5L_info_t *info, void *op)
{
	(void)g;
	(void)info;
	name_list *L = op;
	if (name[0] == 'g') /* "grid" (io.c:176 skips every name starting with 'g') */
		return 0;
	if (L->n < L->max)
		snprintf(L->names[L->n], 64, "%s", name);
	L->n++;
	return 0;
}

static int by_time(const void *a, const void *b)
{
	double x = strtod((const char *)a, NULL), y = strtod((const char *)b, NULL);
	return (x > y) - (x < y);
}

int gray_h5_list_times(const char *path, char (*names)[64], double *times, int max)
{
	hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0)
		return h5fail("%s is not a valid HDF5 file", path);
	H5G_info_t gi;
	if (H5Gget_info(f, &gi) < 0 || gi.nlinks < 2) {
		H5Fclose(f);
		return h5fail("not enough groups in %s", path);
	}
	name_list L = {names, 0, max};
	hsize_t idx = 0;
	herr_t st = H5Literate(f, H5_INDEX_NAME, H5_ITER_NATIVE, &idx, collect, &L);
	H5Fclose(f);
	if (st < 0)
		return h5fail("listing the groups of %s failed", path);
	if (L.n > max)
		return h5fail("more than %d time levels in %s", max, path);
	qsort(names, (size_t)L.n, 64, by_time);
	for (int i = 0; i < L.n; ++i)
		times[i] = strtod(names[i], NULL);
	return L.n;
}

static int read_all(hid_t loc, const char *name, size_t expect, double *out, int *dtype)
{
	hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
	if (ds < 0)
		return h5fail("dataset %s not found", name);
	hid_t sp = H5Dget_space(ds);
	hssize_t np = sp < 0 ? -1 : H5Sget_simple_extent_npoints(sp);
	int rc = 0;
	if (np < 0 || (expect && (size_t)np != expect))
		rc = h5fail("dataset %s has %lld points, expected %zu", name, (long long)np, expect);
	if (!rc && dtype) {
		hid_t t = H5Dget_type(ds);
		*dtype = (int)H5Tget_size(t);
		H5Tclose(t);
	}
	if (!rc && out && H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
		rc = h5fail("reading dataset %s failed", name);
	if (sp >= 0)
		H5Sclose(sp);
	H5Dclose(ds);
	return rc;
}

int gray_h5_read_axes(const char *path, int n[3], double *xs, double *ys, double *zs, int *dtype)
{
	hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0)
		return h5fail("%s is not a valid HDF5 file", path);
	hid_t g = H5Gopen2(f, "grid", H5P_DEFAULT);
	if (g < 0) {
		H5Fclose(f);
		return h5fail("grid group not found in %s", path);
	}
	const char *nm[3] = {"x", "y", "z"};
	double *dst[3] = {xs, ys, zs};
	int rc = 0;
	for (int a = 0; a < 3 && !rc; ++a) {
		hid_t ds = H5Dopen2(g, nm[a], H5P_DEFAULT);
		if (ds < 0) {
			rc = h5fail("grid/%s not found", nm[a]);
			break;
		}
		hid_t sp = H5Dget_space(ds);
		n[a] = (int)H5Sget_simple_extent_npoints(sp);
		H5Sclose(sp);
		H5Dclose(ds);
		if (n[a] < 2)
			rc = h5fail("grid/%s has fewer than 2 points", nm[a]);
		else if (dst[a])
			rc = read_all(g, nm[a], (size_t)n[a], dst[a], a == 0 ? dtype : NULL);
		else if (a == 0 && dtype)
			rc = read_all(g, nm[a], (size_t)n[a], NULL, dtype);
	}
	H5Gclose(g);
	H5Fclose(f);
	return rc;
}

int gray_h5_read_level(const char *path, const char *name, const int n[3], void *gam,
                       void *met, void *rho, int out_dtype)
{
	const size_t np = (size_t)n[0] * n[1] * n[2];
	double *buf = malloc(np * sizeof(double));
	if (!buf)
		return h5fail("out of memory");
	hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0) {
		free(buf);
		return h5fail("%s is not a valid HDF5 file", path);
	}
	hid_t g = H5Gopen2(f, name, H5P_DEFAULT);
	int rc = g < 0 ? h5fail("time %s not found", name) : 0;
	for (int c = 0; !rc && c < 51; ++c) {
		char dn[32];
		void *dst;
		int stride, off;
		if (c < 40) {
			gamma_name(c, dn, sizeof(dn));
			dst = gam; stride = 40; off = c;
		} else if (c < 50) {
			metric_name(c - 40, dn, sizeof(dn));
			dst = met; stride = 10; off = c - 40;
		} else {
			snprintf(dn, sizeof(dn), "rho");
			dst = rho; stride = 1; off = 0;
		}
		/* every dataset must exist with the right size, even if unused
		 * (the reference reader requires all 51, io.c:356-428) */
		rc = read_all(g, dn, np, dst ? buf : NULL, NULL);
		if (rc || !dst)
			continue;
		if (out_dtype == 4)
			for (size_t p = 0; p < np; ++p)
				((float *)dst)[p * stride + off] = (float)buf[p];
		else
			for (size_t p = 0; p < np; ++p)
				((double *)dst)[p * stride + off] = buf[p];
	}
	if (g >= 0)
		H5Gclose(g);
	H5Fclose(f);
	free(buf);
	return rc;
}

int gray_axis_map(const double *x, int n)
{
	if (n < 2)
		return -1;
	const double lo = x[0], hi = x[n - 1], dx = (hi - lo) / (double)(n - 1);
	int uni = 1, fish = 1;
	const double B = cbrt(asinh(lo)), A = (cbrt(asinh(hi)) - B) / (double)(n - 1);
	for (int i = 0; i < n; ++i) {
		if (fabs(x[i] - (lo + (double)i * dx)) > 1e-6 * fabs(dx))
			uni = 0;
		double c = A * (double)i + B, f = sinh(c * c * c);
		if (fabs(x[i] - f) > 1e-6 * (fabs(f) > 1.0 ? fabs(f) : 1.0))
			fish = 0;
	}
	return uni ? 0 : (fish ? 1 : -1);
}

/* The rewrite's ray dump (sim/gray.c:170-176): a new file (H5F_ACC_EXCL,
 * an existing file is an error) with one dataset /rays of the host array
 * palloc(real, n_width, n_height, 8) (sim/gray.c:117), i.e. dims
 * {W, H, nd} with the last index fastest.  `data` holds the states in dump
 * order [H][W][nd] (ray g = gi + gj W); element [gi][gj][c] of the file is
 * data[(gj W + gi) nd + c]. */
int gray_h5_write_rays(const char *path, int w, int h, int nd, const double *data, int dtype)
{
	const size_t n = (size_t)w * h * nd;
	double *buf = malloc(n * sizeof(double));
	if (!buf)
		return h5fail("out of memory");
	for (size_t gi = 0; gi < (size_t)w; ++gi)
		for (size_t gj = 0; gj < (size_t)h; ++gj)
			memcpy(buf + (gi * h + gj) * nd, data + (gj * w + gi) * nd, nd * sizeof(double));
	hid_t f = H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT);
	if (f < 0) {
		free(buf);
		return h5fail("cannot create %s (it must not exist)", path);
	}
	hsize_t dims[3] = {(hsize_t)w, (hsize_t)h, (hsize_t)nd};
	hid_t sp = H5Screate_simple(3, dims, NULL);
	hid_t ds = sp < 0 ? -1 : H5Dcreate2(f, "rays", file_type(dtype), sp, H5P_DEFAULT,
	                                      H5P_DEFAULT, H5P_DEFAULT);
	int rc = ds < 0 || H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0;
	if (ds >= 0)
		H5Dclose(ds);
	if (sp >= 0)
		H5Sclose(sp);
	H5Fclose(f);
	free(buf);
	return rc ? h5fail("writing /rays of %s failed", path) : 0;
}

int gray_h5_read_rays(const char *path, int dims[3], double *out)
{
	hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0)
		return h5fail("cannot open %s", path);
	hid_t ds = H5Dopen2(f, "rays", H5P_DEFAULT);
	int rc = -1;
	if (ds >= 0) {
		hid_t sp = H5Dget_space(ds);
		hsize_t d[3] = {0, 0, 0};
		if (sp >= 0 && H5Sget_simple_extent_ndims(sp) == 3) {
			H5Sget_simple_extent_dims(sp, d, NULL);
			for (int k = 0; k < 3; ++k)
				dims[k] = (int)d[k];
			rc = out ? (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0
			            ? -1 : 0)
			         : 0;
		}
		if (sp >= 0)
			H5Sclose(sp);
		H5Dclose(ds);
	}
	H5Fclose(f);
	return rc ? h5fail("%s has no 3-d /rays dataset", path) : 0;
}
