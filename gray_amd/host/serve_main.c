/*
 * gray-serve -- run the serving request loop from the command line:
 *   gray-serve n=2000 rate=1.0 seed=0x47524159 max_side=1024 slots=1048576 \
 *              budget_gb=8 K=2048 grid_fraction=0.3 verbose=0
 * and print one JSON line with latency percentiles and ray-steps/s.
 */
#define _GNU_SOURCE
#include "../../include/gray_serve.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
	gray_serve_opts o;
	gray_serve_defaults(&o);
	for (int k = 1; k < argc; ++k) {
		char *e = strchr(argv[k], '=');
		if (!e) {
			fprintf(stderr, "gray-serve: expected key=value, got %s\n", argv[k]);
			return 1;
		}
		*e++ = '\0';
		if (!strcmp(argv[k], "n")) o.n_requests = atoi(e);
		else if (!strcmp(argv[k], "rate")) o.rate = atof(e);
		else if (!strcmp(argv[k], "seed")) o.seed = strtoull(e, NULL, 0);
		else if (!strcmp(argv[k], "max_side")) o.max_side = atoi(e);
		else if (!strcmp(argv[k], "slots")) o.slots = strtoull(e, NULL, 0);
		else if (!strcmp(argv[k], "budget_gb")) o.hbm_budget = (size_t)(atof(e) * (1u << 30));
		else if (!strcmp(argv[k], "K")) o.epoch_steps = atoi(e);
		else if (!strcmp(argv[k], "grid_fraction")) o.grid_fraction = atof(e);
		else if (!strcmp(argv[k], "verbose")) o.verbose = atoi(e);
		else if (!strcmp(argv[k], "device")) o.device = atoi(e);
		else {
			fprintf(stderr, "gray-serve: unknown option %s\n", argv[k]);
			return 1;
		}
	}
	if (o.n_requests <= 0 || !(o.rate > 0)) {
		fprintf(stderr, "gray-serve: n and rate must be positive\n");
		return 1;
	}
	gray_serve_report r;
	if (gray_serve_run(&o, &r, NULL))
		return 1;
	printf("{\"requests\": %d, \"wall_s\": %.3f, \"p50_ms\": %.2f, \"p99_ms\": %.2f, "
	       "\"max_ms\": %.2f, \"mean_ms\": %.2f, \"rays\": %lld, \"ray_steps\": [%lld, %lld, %lld], "
	       "\"epochs\": %lld, \"ray_steps_per_s\": %.6e}\n",
	       r.n_done, r.wall_s, r.p50_ms, r.p99_ms, r.max_ms, r.mean_ms, (long long)r.rays,
	       (long long)r.ray_steps[0], (long long)r.ray_steps[1], (long long)r.ray_steps[2],
	       (long long)r.epochs, r.ray_steps_per_s);
	return 0;
}
