/*
 * gray_grid.c -- numerical spacetimes for the gray job (coordinates=dyst):
 * the HDF5 checks, snapshot selection and sliding two-level time window of
 * sim-org/gray.c:30-59 and :147-306, over a resident device spacetime
 * (grayhip_spacetime) and a gridded render (grayhip_ctx).
 *
 * The next level of a sliding window is read from the file on a host thread
 * while the GPU integrates the interval before the slide (the reference
 * reads it afterwards, gray.c:271-289).
 *
 * Decisions for behaviour the reference leaves undefined:
 *   - rays start at q^t = 0 while the grid time is t_init + q^t, and the
 *     time weight is clamped, so only t_init must lie inside the data
 *     (gray.c:194-204 also requires t_final, which with dt_dump < 0 and
 *     levels {0, 1} admits only frozen runs);
 *   - a frozen run uses the level bracketing t_init (gray.c:214 means to);
 *   - the coordinate map is detected from the axes (uniform or fisheye).
 */
#define _GNU_SOURCE
#include "gray_grid.h"
#include "../../include/gray_spacetime.h"

#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define MAX_TIMES 1024

struct gray_grid {
	grayhip_config cfg;
	grayhip_spacetime *st;
	grayhip_ctx *ctx;
	char path[4096];
	char (*names)[64];
	double *times;
	int n_times;
	int n[3];
	int dtype;
	size_t snap;         /* level in slot 0 */
	int frozen;
	void *gam, *met;     /* host staging of one level */
	void *rho;           /* and its density (transfer only) */
	/* prefetch: the level the window needs next is read from the file on a
	 * host thread while the GPU integrates the current interval */
	pthread_t pf_thread;
	int pf_active;       /* a thread is running or has finished unjoined */
	size_t pf_level;
	void *pf_gam, *pf_met, *pf_rho;
	int pf_rc;
	char pf_err[256];
	char err[512];
};

static char g_err[512];

static int gfail(gray_grid *g, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g ? g->err : g_err, 512, fmt, ap);
	va_end(ap);
	return -1;
}

const char *gray_grid_error(const gray_grid *g) { return g ? g->err : g_err; }

int gray_grid_create(const grayhip_config *cfg, gray_grid **out)
{
	*out = NULL;
	if (cfg->n_freq > 0 && cfg->fluid != GRAYHIP_FLUID_THERMAL)
		return gfail(NULL, "radiative transfer on a numerical spacetime uses the reference "
		                   "flow (fluid=thermal)");
	gray_grid *g = calloc(1, sizeof(gray_grid));
	if (!g)
		return gfail(NULL, "out of memory");
	g->cfg = *cfg;
	*out = g;
	return 0;
}

void gray_grid_destroy(gray_grid *g)
{
	if (!g)
		return;
	/* a prefetch thread still reading uses names, path and the pf_
	 * buffers: it ends before anything is freed */
	if (g->pf_active) {
		pthread_join(g->pf_thread, NULL);
		g->pf_active = 0;
	}
	if (g->ctx)
		grayhip_destroy(g->ctx);
	if (g->st)
		grayhip_spacetime_destroy(g->st);
	free(g->names);
	free(g->times);
	free(g->gam);
	free(g->met);
	free(g->rho);
	free(g->pf_gam);
	free(g->pf_met);
	free(g->pf_rho);
	free(g);
}

/* sim-org/gray.c:41-59: the level l with times[l] <= t <= times[l+1]
 * (clamped to the data). */
static size_t find_snapshot(const gray_grid *g, double t)
{
	if (g->n_times < 2 || t <= g->times[0])
		return 0;
	for (int l = 0; l + 1 < g->n_times; ++l)
		if (t >= g->times[l] && t <= g->times[l + 1])
			return (size_t)l;
	return (size_t)(g->n_times - 2);
}

static void *prefetch_main(void *arg)
{
	gray_grid *g = arg;
	g->pf_rc = gray_h5_read_level(g->path, g->names[g->pf_level], g->n, g->pf_gam, g->pf_met,
	                              g->pf_rho, g->dtype);
	if (g->pf_rc)
		snprintf(g->pf_err, sizeof(g->pf_err), "%s", gray_h5_error());
	return NULL;
}

/* Start reading `level` into the prefetch buffers (no-op if one is in
 * flight).  The main thread makes no HDF5 call until it is joined. */
static void prefetch_start(gray_grid *g, size_t level)
{
	if (g->pf_active || level >= (size_t)g->n_times)
		return;
	const size_t np = (size_t)g->n[0] * g->n[1] * g->n[2], sz = (size_t)g->dtype;
	if (!g->pf_gam && !(g->pf_gam = malloc(np * 40 * sz)))
		return;
	if (!g->pf_met && !(g->pf_met = malloc(np * 10 * sz)))
		return;
	if (g->rho && !g->pf_rho && !(g->pf_rho = malloc(np * sz)))
		return;
	g->pf_level = level;
	g->pf_rc = 0;
	if (pthread_create(&g->pf_thread, NULL, prefetch_main, g) == 0)
		g->pf_active = 1;
}

static int upload(gray_grid *g, size_t level, int slot, const void *gam, const void *met,
                  const void *rho)
{
	if (grayhip_spacetime_upload(g->st, slot, g->times[level], gam, met))
		return gfail(g, "upload of level %s: %s", g->names[level], grayhip_spacetime_error(g->st));
	if (rho && grayhip_spacetime_upload_rho(g->st, slot, rho))
		return gfail(g, "density of level %s: %s", g->names[level],
		             grayhip_spacetime_error(g->st));
	return 0;
}

static int load_level(gray_grid *g, size_t level, int slot)
{
	if (g->pf_active) {
		pthread_join(g->pf_thread, NULL);
		g->pf_active = 0;
		if (g->pf_level == level) {
			if (g->pf_rc)
				return gfail(g, "%s", g->pf_err);
			return upload(g, level, slot, g->pf_gam, g->pf_met, g->rho ? g->pf_rho : NULL);
		}
	}
	if (gray_h5_read_level(g->path, g->names[level], g->n, g->gam, g->met, g->rho, g->dtype))
		return gfail(g, "%s", gray_h5_error());
	return upload(g, level, slot, g->gam, g->met, g->rho);
}

int gray_grid_load(gray_grid *g, const char *path, double t_init, double dt_dump, size_t i_init,
                   size_t n_dump, size_t fast_light)
{
	(void)dt_dump;
	(void)i_init;
	(void)n_dump;
	printf("GRay2: Reading spacetime from file %s\n", path);
	if (access(path, F_OK) != 0)
		return gfail(g, "File %s could not be read", path);
	snprintf(g->path, sizeof(g->path), "%s", path);
	g->names = calloc(MAX_TIMES, 64);
	g->times = calloc(MAX_TIMES, sizeof(double));
	if (!g->names || !g->times)
		return gfail(g, "out of memory");
	g->n_times = gray_h5_list_times(path, g->names, g->times, MAX_TIMES);
	if (g->n_times < 0)
		return gfail(g, "%s", gray_h5_error());
	if (g->n_times < 1)
		return gfail(g, "no time level in %s", path);

	double *ax[3];
	int dt;
	if (gray_h5_read_axes(path, g->n, NULL, NULL, NULL, &dt))
		return gfail(g, "%s", gray_h5_error());
	for (int a = 0; a < 3; ++a)
		ax[a] = malloc((size_t)g->n[a] * sizeof(double));
	int map = -2, rc = 0;
	if (!ax[0] || !ax[1] || !ax[2])
		rc = gfail(g, "out of memory");
	else if (gray_h5_read_axes(path, g->n, ax[0], ax[1], ax[2], &dt))
		rc = gfail(g, "%s", gray_h5_error());
	grayhip_grid_desc d;
	memset(&d, 0, sizeof(d));
	for (int a = 0; a < 3 && !rc; ++a) {
		int m = gray_axis_map(ax[a], g->n[a]);
		if (m < 0 || (map >= 0 && m != map))
			rc = gfail(g, "grid axis %c is neither uniform nor fisheye (sinh((A i + B)^3))",
			           "xyz"[a]);
		map = m;
		d.n[a] = g->n[a];
		d.lo[a] = ax[a][0];
		d.hi[a] = ax[a][g->n[a] - 1];
	}
	for (int a = 0; a < 3; ++a)
		free(ax[a]);
	if (rc)
		return rc;
	d.map = map == 1 ? GRAYHIP_MAP_FISHEYE : GRAYHIP_MAP_UNIFORM;
	d.dtype = g->dtype = (dt == 8 ? 8 : 4);
	d.device = g->cfg.device;

	const size_t np = (size_t)g->n[0] * g->n[1] * g->n[2];
	g->gam = malloc(np * 40 * (size_t)g->dtype);
	g->met = malloc(np * 10 * (size_t)g->dtype);
	if (!g->gam || !g->met)
		return gfail(g, "out of memory");
	if (g->cfg.n_freq > 0 && !(g->rho = malloc(np * (size_t)g->dtype)))
		return gfail(g, "out of memory");
	if (grayhip_spacetime_create(&d, &g->st))
		return gfail(g, "cannot allocate the spacetime on the device");

	/* gray.c:185-241 */
	g->frozen = fast_light != 0;
	if (g->n_times == 1) {
		printf("Found only one time in data, freezing spacetime\n");
		g->frozen = 1;
	} else if (t_init < g->times[0] || t_init > g->times[g->n_times - 1]) {
		return gfail(g, "t_init (%4.1f) is outside domain of the data (%5.1f, %5.1f)", t_init,
		             g->times[0], g->times[g->n_times - 1]);
	}
	g->snap = find_snapshot(g, t_init);
	if (load_level(g, g->snap, 0))
		return -1;
	if (g->frozen) {
		printf("Assuming fast light\n");
		grayhip_spacetime_set_frozen(g->st, 1);
	} else {
		printf("Working with slow light\n");
		if (load_level(g, g->snap + 1, 1))
			return -1;
		grayhip_spacetime_set_frozen(g->st, 0);
	}

	grayhip_config cfg = g->cfg;
	cfg.spacetime = g->st;
	cfg.t_init = t_init;
	if (grayhip_create(&cfg, &g->ctx))
		return gfail(g, "grayhip_create: %s", grayhip_global_error());
	return 0;
}

int gray_grid_icond(gray_grid *g, double w_img, double h_img, double r_obs, double i_obs,
                    double j_obs)
{
	if (!g->ctx)
		return gfail(g, "spacetime not loaded");
	if (grayhip_icond(g->ctx, w_img, h_img, r_obs, i_obs, j_obs))
		return gfail(g, "%s", grayhip_last_error(g->ctx));
	return 0;
}

int gray_grid_evolve(gray_grid *g, double t, double target, size_t n_sub, double *ns)
{
	if (!g->ctx)
		return gfail(g, "spacetime not loaded");
	const double t1 = g->times[g->snap];
	const double t2 = g->n_times > 1 ? g->times[g->snap + 1] : t1;
	const int slide = !g->frozen && (target < t1 || target > t2);
	if (slide) {
		/* read the level the slide below will need while the GPU runs */
		size_t s = find_snapshot(g, target);
		if (s == g->snap + 1)
			prefetch_start(g, s + 1);
		else if (s != g->snap)
			prefetch_start(g, s);
	}
	if (grayhip_evolve(g->ctx, target - t, n_sub, ns))
		return gfail(g, "%s", grayhip_last_error(g->ctx));
	/* gray.c:259-293: slide the window when target leaves it */
	if (slide) {
		size_t old = g->snap;
		size_t s = find_snapshot(g, target);
		if (s == old)
			return 0; /* clamped at an end of the data */
		g->snap = s;
		if (s == old + 1) {
			grayhip_spacetime_shift(g->st, 1);
			if (load_level(g, s + 1, 1))
				return -1;
		} else if (s + 1 == old) {
			grayhip_spacetime_shift(g->st, 0);
			if (load_level(g, s, 0))
				return -1;
		} else if (load_level(g, s, 0) || load_level(g, s + 1, 1)) {
			return -1;
		}
		grayhip_spacetime_set_frozen(g->st, 0);
	}
	return 0;
}

int gray_grid_read(gray_grid *g, void *data, void *info)
{
	if (!g->ctx)
		return gfail(g, "spacetime not loaded");
	if (grayhip_read(g->ctx, data, info))
		return gfail(g, "%s", grayhip_last_error(g->ctx));
	return 0;
}
