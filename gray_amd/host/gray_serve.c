/*
 * gray_serve.c -- the C request loop of the serving form: a seeded Poisson
 * stream of uneven render requests submitted to one grayhip_server pool as
 * they arrive; the pool admits them by HBM budget and refills ray slots as
 * rays finish, so requests overlap on the device.  Reports latency
 * percentiles and ray-steps/s.
 */
#define _GNU_SOURCE
#include "../../include/gray_serve.h"
#include "../../include/gray_spacetime.h"
#include "../../include/grayhip.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static uint64_t sm64(uint64_t *s)
{
	uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

static double unif(uint64_t *s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void gray_serve_defaults(gray_serve_opts *o)
{
	memset(o, 0, sizeof(*o));
	o->n_requests = 2000;
	o->seed = 0x47524159ull;
	o->rate = 1.0;
	o->grid_fraction = 0.3;
	o->max_side = 1024;
	o->grid_n[0] = 96;
	o->grid_n[1] = 128;
	o->slots = 1u << 20;
	o->hbm_budget = (size_t)8 << 30;
	o->epoch_steps = 2048;
}

int gray_serve_stream(const gray_serve_opts *o, gray_serve_item *it)
{
	static const int sides[5] = {64, 128, 256, 512, 1024};
	static const double w[5] = {0.3, 0.3, 0.2, 0.15, 0.05};
	uint64_t s = o->seed;
	double t = 0.0;
	for (int k = 0; k < o->n_requests; ++k) {
		t += -log(1.0 - unif(&s)) / o->rate;
		double u = unif(&s), c = 0.0;
		int side = sides[4];
		for (int j = 0; j < 5; ++j) {
			c += w[j];
			if (u < c) {
				side = sides[j];
				break;
			}
		}
		if (side > o->max_side)
			side = o->max_side;
		it[k].side = side;
		it[k].a_spin = 0.998 * unif(&s);
		it[k].i_obs = acos(unif(&s)) * (180.0 / M_PI);
		it[k].n_freq = 1 + (int)(4.0 * unif(&s));
		if (it[k].n_freq > 4)
			it[k].n_freq = 4;
		it[k].grid = unif(&s) < o->grid_fraction ? (unif(&s) < 0.5 ? 0 : 1) : -1;
		it[k].arrival_s = t;
	}
	return 0;
}

static void fill_request(const gray_serve_item *it, grayhip_spacetime *const grids[2],
                         grayhip_request *q)
{
	static const double nus[4] = {1.0, 2.67, 4.0, 8.0};
	/* radio to sub-mm frequencies (Hz) for the thermal emission model */
	static const double grid_nus[4] = {2.3e11, 3.45e11, 5.0e11, 8.6e11};
	memset(q, 0, sizeof(*q));
	q->w_rays = q->h_rays = (size_t)it->side;
	q->i_obs = it->i_obs;
	q->escape = GRAYHIP_ESC_OUTWARD;
	q->n_sub = 1024;
	q->dt = -32.0;
	if (it->grid >= 0) {
		/* camera at r = 35 M inside the 80 M box, 4 dumps of -32; the
		 * channels through the reference flow on the grid's density */
		q->spacetime = grids[it->grid];
		q->w_img = q->h_img = 32.0;
		q->r_obs = 35.0;
		q->n_dump = 4;
		q->t_init = 1.0;
		q->n_freq = it->n_freq;
		for (int i = 0; i < it->n_freq; ++i)
			q->nu[i] = grid_nus[i];
		q->fluid = GRAYHIP_FLUID_THERMAL;
	} else {
		/* camera at r = 1000 M, synthetic disk, 128 dumps of -32 */
		q->a_spin = it->a_spin;
		q->w_img = q->h_img = 64.0;
		q->r_obs = 1000.0;
		q->r_esc = 1000.0;
		q->n_dump = 128;
		q->n_freq = it->n_freq;
		for (int i = 0; i < it->n_freq; ++i)
			q->nu[i] = nus[i];
		q->fluid = GRAYHIP_FLUID_DISK;
	}
}

static int make_grid(int n, int device, grayhip_spacetime **out)
{
	const size_t np = (size_t)n * n * n;
	double *xs = malloc((size_t)n * sizeof(double));
	double *gam = malloc(np * 40 * sizeof(double)), *met = malloc(np * 10 * sizeof(double));
	double *rho = malloc(np * sizeof(double));
	float *g32 = malloc(np * 40 * sizeof(float)), *m32 = malloc(np * 10 * sizeof(float));
	float *r32 = malloc(np * sizeof(float));
	int rc = -1;
	if (!xs || !gam || !met || !rho || !g32 || !m32 || !r32)
		goto done;
	gray_axis_uniform(-40.0, 40.0, n, xs);
	grayhip_grid_desc d;
	memset(&d, 0, sizeof(d));
	for (int a = 0; a < 3; ++a) {
		d.n[a] = n;
		d.lo[a] = -40.0;
		d.hi[a] = 40.0;
	}
	d.map = GRAYHIP_MAP_UNIFORM;
	d.dtype = 4;
	d.device = device;
	if (grayhip_spacetime_create(&d, out))
		goto done;
	gray_bks b;
	const double ax[3] = {1.0, 0.0, 0.0};
	gray_bks_setup(&b, 1.0, 0.7, 0.3, ax);
	for (int l = 0; l < 2; ++l) {
		gray_bks_sample(&b, (double)l, n, n, n, xs, xs, xs, gam, met, rho);
		for (size_t i = 0; i < np * 40; ++i)
			g32[i] = (float)gam[i];
		for (size_t i = 0; i < np * 10; ++i)
			m32[i] = (float)met[i];
		for (size_t i = 0; i < np; ++i)
			r32[i] = (float)rho[i];
		if (grayhip_spacetime_upload(*out, l, (double)l, g32, m32) ||
		    grayhip_spacetime_upload_rho(*out, l, r32))
			goto done;
	}
	rc = grayhip_spacetime_set_frozen(*out, 0);
done:
	free(xs);
	free(gam);
	free(met);
	free(rho);
	free(g32);
	free(m32);
	free(r32);
	return rc;
}

static int cmp_d(const void *a, const void *b)
{
	double x = *(const double *)a, y = *(const double *)b;
	return (x > y) - (x < y);
}

int gray_serve_run_items(const gray_serve_opts *o, const gray_serve_item *it, int n,
                         gray_serve_report *rep, double *latency_ms, double *const *results)
{
	double *t_done = calloc((size_t)(n > 0 ? n : 1), sizeof(double));
	int64_t *ids = calloc((size_t)(n > 0 ? n : 1), sizeof(int64_t));
	grayhip_spacetime *grids[2] = {NULL, NULL};
	grayhip_server *srv = NULL;
	int rc = -1;
	memset(rep, 0, sizeof(*rep));
	if (n <= 0 || !t_done || !ids)
		goto out;
	for (int k = 1; k < n; ++k)
		if (it[k].arrival_s < it[k - 1].arrival_s) {
			fprintf(stderr, "gray-serve: arrivals must be in time order\n");
			goto out;
		}
	for (int g = 0; g < 2; ++g) {
		int used = 0;
		for (int k = 0; k < n; ++k)
			used |= it[k].grid == g;
		if (used && make_grid(o->grid_n[g], o->device, &grids[g])) {
			fprintf(stderr, "gray-serve: cannot build grid %d\n", g);
			goto out;
		}
	}
	grayhip_server_config sc = {o->device, o->slots, o->hbm_budget, o->epoch_steps, 4096};
	if (grayhip_server_create(&sc, &srv)) {
		fprintf(stderr, "gray-serve: %s\n", grayhip_server_error(NULL));
		goto out;
	}
	int next = 0, n_done = 0;
	const double t0 = now_s() - it[0].arrival_s;
	while (n_done < n) {
		double t = now_s() - t0;
		for (; next < n && it[next].arrival_s <= t; ++next) {
			grayhip_request q;
			fill_request(&it[next], grids, &q);
			if (grayhip_server_submit(srv, &q, &ids[next])) {
				fprintf(stderr, "gray-serve: submit: %s\n", grayhip_server_error(srv));
				goto out;
			}
		}
		int busy = 0;
		if (grayhip_server_step(srv, &busy)) {
			fprintf(stderr, "gray-serve: step: %s\n", grayhip_server_error(srv));
			goto out;
		}
		int64_t id;
		while (grayhip_server_next_done(srv, &id) == 0) {
			double td = now_s() - t0;
			/* ids are issued in submission (= list) order */
			int k = (int)(id - ids[0]);
			double *res = NULL;
			if (k >= 0 && k < n) {
				t_done[k] = td;
				rep->rays += (int64_t)it[k].side * it[k].side;
				res = results ? results[k] : NULL;
			}
			if (grayhip_server_result(srv, id, res, NULL, NULL)) {
				fprintf(stderr, "gray-serve: result: %s\n", grayhip_server_error(srv));
				goto out;
			}
			++n_done;
			if (o->verbose) {
				printf("request %d (%dx%d %s) done: %.1f ms\n", k, it[k].side, it[k].side,
				       it[k].grid >= 0 ? "grid" : "kerr",
				       1e3 * (td - it[k].arrival_s));
				fflush(stdout);
			}
		}
		if (!busy && next < n) {
			double wait = it[next].arrival_s - (now_s() - t0);
			if (wait > 0) {
				struct timespec ts = {(time_t)wait, (long)((wait - floor(wait)) * 1e9)};
				nanosleep(&ts, NULL);
			}
		}
	}
	double *lat = malloc((size_t)n * sizeof(double));
	if (!lat)
		goto out;
	double sum = 0.0, last = 0.0;
	for (int k = 0; k < n; ++k) {
		lat[k] = 1e3 * (t_done[k] - it[k].arrival_s);
		sum += lat[k];
		if (t_done[k] > last)
			last = t_done[k];
		if (latency_ms)
			latency_ms[k] = lat[k];
	}
	qsort(lat, (size_t)n, sizeof(double), cmp_d);
	rep->n_done = n;
	rep->p50_ms = lat[(n - 1) / 2];
	rep->p99_ms = lat[(int)ceil(0.99 * n) - 1];
	rep->max_ms = lat[n - 1];
	rep->mean_ms = sum / n;
	rep->wall_s = last - it[0].arrival_s;
	grayhip_server_stats(srv, rep->ray_steps, &rep->epochs);
	rep->ray_steps_per_s =
	    (double)(rep->ray_steps[0] + rep->ray_steps[1] + rep->ray_steps[2]) / rep->wall_s;
	free(lat);
	rc = 0;
out:
	grayhip_server_destroy(srv);
	grayhip_spacetime_destroy(grids[0]);
	grayhip_spacetime_destroy(grids[1]);
	free(t_done);
	free(ids);
	return rc;
}

int gray_serve_run(const gray_serve_opts *o, gray_serve_report *rep, double *latency_ms)
{
	gray_serve_item *it = calloc((size_t)(o->n_requests > 0 ? o->n_requests : 1), sizeof(*it));
	if (!it)
		return -1;
	gray_serve_stream(o, it);
	int rc = gray_serve_run_items(o, it, o->n_requests, rep, latency_ms, NULL);
	free(it);
	return rc;
}
