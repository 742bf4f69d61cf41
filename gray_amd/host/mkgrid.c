/*
 * gray-mkgrid -- write a boosted Kerr-Schild numerical spacetime in the
 * reference's HDF5 layout (the role of tools/generate_data.py):
 *
 *   gray-mkgrid out=data.h5 n=128 lo=-40 hi=40 M=1 a=0.7 v=0.3 axis=x \
 *               times=0.0,1.0 dtype=4 fisheye=0
 *
 * Writes /grid/{x,y,z} and one group per time with 40 Gamma_*, 10 g_* and
 * rho (= 1/|x|, as generate_data.py:124-128), in FP32 (dtype=4) or FP64.
 */
#define _GNU_SOURCE
#include "../../include/gray_spacetime.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
	const char *out = "data.h5";
	int n = 128, dtype = 4, fisheye = 0;
	double lo = -40, hi = 40, m = 1, a = 0.7, v = 0.3, ax[3] = {1, 0, 0};
	char times[1024] = "0.0,1.0";
	for (int k = 1; k < argc; ++k) {
		char *e = strchr(argv[k], '=');
		if (!e) {
			fprintf(stderr, "gray-mkgrid: expected key=value, got %s\n", argv[k]);
			return 1;
		}
		*e++ = '\0';
		if (!strcmp(argv[k], "out")) out = e;
		else if (!strcmp(argv[k], "n")) n = atoi(e);
		else if (!strcmp(argv[k], "lo")) lo = atof(e);
		else if (!strcmp(argv[k], "hi")) hi = atof(e);
		else if (!strcmp(argv[k], "M")) m = atof(e);
		else if (!strcmp(argv[k], "a")) a = atof(e);
		else if (!strcmp(argv[k], "v")) v = atof(e);
		else if (!strcmp(argv[k], "dtype")) dtype = atoi(e);
		else if (!strcmp(argv[k], "fisheye")) fisheye = atoi(e);
		else if (!strcmp(argv[k], "times")) snprintf(times, sizeof(times), "%s", e);
		else if (!strcmp(argv[k], "axis")) {
			ax[0] = ax[1] = ax[2] = 0;
			ax[e[0] == 'x' ? 0 : (e[0] == 'y' ? 1 : 2)] = 1;
		} else {
			fprintf(stderr, "gray-mkgrid: unknown option %s\n", argv[k]);
			return 1;
		}
	}
	if (n < 2 || (dtype != 4 && dtype != 8) || !(v > -1 && v < 1)) {
		fprintf(stderr, "gray-mkgrid: invalid n, dtype or v\n");
		return 1;
	}
	double *xs = malloc((size_t)n * sizeof(double));
	size_t np = (size_t)n * n * n;
	double *gam = malloc(np * 40 * sizeof(double)), *met = malloc(np * 10 * sizeof(double));
	double *rho = malloc(np * sizeof(double));
	if (!xs || !gam || !met || !rho) {
		fprintf(stderr, "gray-mkgrid: out of memory\n");
		return 1;
	}
	(fisheye ? gray_axis_fisheye : gray_axis_uniform)(lo, hi, n, xs);
	if (gray_h5_create(out, n, n, n, xs, xs, xs, dtype)) {
		fprintf(stderr, "gray-mkgrid: %s\n", gray_h5_error());
		return 1;
	}
	gray_bks b;
	gray_bks_setup(&b, m, a, v, ax);
	for (char *t = strtok(times, ","); t; t = strtok(NULL, ",")) {
		gray_bks_sample(&b, atof(t), n, n, n, xs, xs, xs, gam, met, rho);
		if (gray_h5_write_level(out, t, n, n, n, gam, met, rho, dtype)) {
			fprintf(stderr, "gray-mkgrid: %s\n", gray_h5_error());
			return 1;
		}
		printf("gray-mkgrid: wrote level %s\n", t);
	}
	free(xs); free(gam); free(met); free(rho);
	return 0;
}
