/*
 * gray_io.c -- raw dump files in the layout of sim-org/io.c:39-72:
 * four size_t (sz, n_data, w_rays, h_rays), n_rays*n_data reals, then
 * n_rays*n_info reals of diagnostics (n_info = 1).  tools/gray.py:22-43
 * reads the same layout.
 */
#include "../../include/gray_job.h"

#include <stdio.h>
#include <stdlib.h>

int gray_write_raw(const char *path, size_t sz, size_t n_data, size_t w_rays,
                   size_t h_rays, const void *data, const void *info)
{
	const size_t n_rays = w_rays * h_rays;
	FILE *f = fopen(path, "wb");
	if (!f)
		return -1;
	size_t hdr[4] = {sz, n_data, w_rays, h_rays};
	int ok = fwrite(hdr, sizeof(size_t), 4, f) == 4 &&
	         fwrite(data, sz * n_data, n_rays, f) == n_rays &&
	         fwrite(info, sz, n_rays, f) == n_rays;
	if (fclose(f) != 0)
		ok = 0;
	return ok ? 0 : -1;
}

int gray_read_raw(const char *path, size_t hdr[4], void *data, void *info)
{
	FILE *f = fopen(path, "rb");
	if (!f)
		return -1;
	int ok = fread(hdr, sizeof(size_t), 4, f) == 4 && (hdr[0] == 4 || hdr[0] == 8);
	if (ok) {
		const size_t n_rays = hdr[2] * hdr[3];
		if (data)
			ok = fread(data, hdr[0] * hdr[1], n_rays, f) == n_rays;
		else
			ok = fseek(f, (long)(hdr[0] * hdr[1] * n_rays), SEEK_CUR) == 0;
		if (ok && info)
			ok = fread(info, hdr[0], n_rays, f) == n_rays;
	}
	fclose(f);
	return ok ? 0 : -1;
}
