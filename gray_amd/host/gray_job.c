/*
 * gray_job.c -- the gray job module: conf / init / exec over the gfx950
 * device layer (include/grayhip.h).
 *
 * Life cycle and printed progress follow sim-org/gray.c:62-346:
 *   conf  -- parse one key=value into the option tables (gray.c:62-111);
 *   init  -- allocate the device state for W*H rays (gray.c:113-145);
 *   exec  -- initial conditions, dump 0, then n_dump evolve+dump rounds of
 *            dt_dump each (gray.c:147-306), printing ns/step/ray.
 */
#define _GNU_SOURCE
#include "../../include/gray_job.h"
#include "../../include/gray_spacetime.h"
#include "../../include/grayhip.h"
#include "gray_grid.h"
#include "gray_opts.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

struct gray {
	gray_job_ops super;            /* first member: the job methods */
	struct gray_options o;
	size_t n_coor, n_freq, n_info; /* 8, len(nu), 1 (gray.c:323-325) */
	grayhip_ctx *ctx;
	gray_grid *grid;               /* numerical spacetime (coordinates=dyst) */
	void *data, *info, *tmp;       /* host staging for dumps */
	int verbose;
};

#define EGO ((struct gray *)ego)

/* The rewrite's job surface (sim/gray.c): selected by spacetime=..., its
 * planner switch (sim/gray.c:49-58), with initcond=infcam and ray files. */
static int rewrite(const struct gray_options *o)
{
	return o->spacetime && *o->spacetime;
}

static int analytic(const struct gray_options *o)
{
	return (o->coordinates && !strcmp(o->coordinates, "KS")) || rewrite(o);
}

static int _conf(void *ego, const char *arg)
{
	char err[256];
	double *old_nu = EGO->o.nu;
	int rc = gray_options_conf(&EGO->o, arg, err, sizeof(err));
	if (rc) {
		fprintf(stderr, "GRay2: %s\n", err);
		return rc;
	}
	if (EGO->o.nu != old_nu) { /* nu was configured (gray.c:92-108) */
		EGO->n_freq = gray_options_n_freq(&EGO->o);
		printf("nu: [");
		for (size_t i = 0; i < EGO->n_freq; ++i)
			printf(i ? ", %f" : "%f", EGO->o.nu[i]);
		printf("]\n");
	}
	return 0;
}

static int pick_fluid(const char *s, int analytic, int *out)
{
	if (!s || !*s)
		*out = analytic ? GRAYHIP_FLUID_DISK : GRAYHIP_FLUID_THERMAL;
	else if (!strcmp(s, "disk"))
		*out = GRAYHIP_FLUID_DISK;
	else if (!strcmp(s, "thermal"))
		*out = GRAYHIP_FLUID_THERMAL;
	else if (!strcmp(s, "uniform"))
		*out = GRAYHIP_FLUID_UNIFORM;
	else if (!strcmp(s, "none"))
		*out = GRAYHIP_FLUID_NONE;
	else
		return -1;
	return 0;
}

static int _init(void *ego)
{
	struct gray_options *o = &EGO->o;
	const size_t n_rays = o->w_rays * o->h_rays;
	const size_t n_data = EGO->n_coor + 2 * EGO->n_freq;
	const size_t sz = o->precision;

	if (!o->scheme || strcmp(o->scheme, "RK4")) {
		fprintf(stderr, "GRay2: unknown integration scheme \"%s\" (RK4 only)\n", o->scheme);
		return EXIT_FAILURE;
	}
	if (rewrite(o)) {
		/* sim/gray.c:49-70: Kerr and infcam are the only modules; the rays
		 * are real[8] (sim/gray.c:116-117), so no transfer channels */
		if (strcmp(o->spacetime, "Kerr")) {
			fprintf(stderr, "GRay2: Unknown spacetime configuration \"%s\"\n", o->spacetime);
			return EXIT_FAILURE;
		}
		if (!o->initcond || strcmp(o->initcond, "infcam")) {
			fprintf(stderr, "GRay2: Unknown initial conditions for rays \"%s\"\n",
			        o->initcond ? o->initcond : "");
			return EXIT_FAILURE;
		}
		if (o->coordinates && strcmp(o->coordinates, "pxcenter") &&
		    strcmp(o->coordinates, "dyst") && strcmp(o->coordinates, "KS")) {
			fprintf(stderr, "GRay2: infcam coordinates \"%s\" (pxcenter only)\n",
			        o->coordinates);
			return EXIT_FAILURE;
		}
		if (EGO->n_freq) {
			fprintf(stderr, "GRay2: the ray files of spacetime=Kerr hold 8 reals per ray "
			                "(no nu)\n");
			return EXIT_FAILURE;
		}
	}
	if (!o->morder || (strcmp(o->morder, "AoS") && strcmp(o->morder, "SoA"))) {
		fprintf(stderr, "GRay2: unknown memory order \"%s\" (AoS or SoA)\n", o->morder);
		return EXIT_FAILURE;
	}
	if (n_rays == 0) {
		fprintf(stderr, "GRay2: w_rays and h_rays must be positive\n");
		return EXIT_FAILURE;
	}
	if (EGO->n_freq > GRAYHIP_MAX_FREQ) {
		fprintf(stderr, "GRay2: at most %d frequency channels\n", GRAYHIP_MAX_FREQ);
		return EXIT_FAILURE;
	}

	grayhip_config cfg;
	memset(&cfg, 0, sizeof(cfg));
	cfg.w_rays = o->w_rays;
	cfg.h_rays = o->h_rays;
	cfg.precision = (int)sz;
	cfg.a_spin = o->a_spin;
	cfg.n_freq = (int)EGO->n_freq;
	for (size_t i = 0; i < EGO->n_freq; ++i)
		cfg.nu[i] = o->nu[i];
	if (pick_fluid(o->fluid, analytic(o), &cfg.fluid)) {
		fprintf(stderr, "GRay2: unknown fluid \"%s\" (disk, uniform, thermal, none)\n",
		        o->fluid);
		return EXIT_FAILURE;
	}
	cfg.m_adm = o->M_ADM;
	cfg.b_field = o->b_field;
	if (!EGO->n_freq)
		cfg.fluid = GRAYHIP_FLUID_NONE;
	if (!o->escape || !strcmp(o->escape, "outward"))
		cfg.escape = GRAYHIP_ESC_OUTWARD;
	else if (!strcmp(o->escape, "ref"))
		cfg.escape = GRAYHIP_ESC_REF;
	else {
		fprintf(stderr, "GRay2: unknown escape rule \"%s\" (outward, ref)\n", o->escape);
		return EXIT_FAILURE;
	}
	cfg.r_esc = o->r_esc;
	cfg.eps_stop = o->eps_stop;
	cfg.device = (int)o->i_device;

	if (!analytic(o)) {
		/* numerical spacetime: the grid is loaded by exec (gray.c:171-241) */
		int rc = gray_grid_create(&cfg, &EGO->grid);
		if (rc) {
			fprintf(stderr, "GRay2: %s\n", gray_grid_error(NULL));
			return EXIT_FAILURE;
		}
	} else {
		int rc = grayhip_create(&cfg, &EGO->ctx);
		if (rc) {
			fprintf(stderr, "GRay2: grayhip_create: %s\n", grayhip_global_error());
			return EXIT_FAILURE;
		}
	}
	EGO->data = malloc(n_rays * n_data * (rewrite(o) ? sizeof(double) : sz));
	EGO->info = malloc(n_rays * sz);
	EGO->tmp = malloc(n_rays * n_data * sz);
	if (!EGO->data || !EGO->info || !EGO->tmp) {
		fprintf(stderr, "GRay2: host allocation failed\n");
		return EXIT_FAILURE;
	}
	return EXIT_SUCCESS;
}

/* sim-org/io.c:39-72: write the device state (in the configured memory
 * order) and the null-constraint diagnostic to outfile % i. */
static int dump(void *ego, size_t i)
{
	struct gray_options *o = &EGO->o;
	const size_t n_rays = o->w_rays * o->h_rays;
	const size_t n_data = EGO->n_coor + 2 * EGO->n_freq;
	const size_t sz = o->precision;
	int rc = EGO->ctx ? grayhip_read(EGO->ctx, EGO->data, EGO->info)
	                  : gray_grid_read(EGO->grid, EGO->data, EGO->info);
	if (rc) {
		fprintf(stderr, "GRay2: reading the ray states failed: %s\n",
		        EGO->ctx ? grayhip_last_error(EGO->ctx) : gray_grid_error(EGO->grid));
		return -1;
	}
	const void *out = EGO->data;
	if (!strcmp(o->morder, "SoA")) { /* device order of sim-org/SoA.cl:29 */
		char *dst = EGO->tmp;
		const char *src = EGO->data;
		for (size_t g = 0; g < n_rays; ++g)
			for (size_t s = 0; s < n_data; ++s)
				memcpy(dst + (s * n_rays + g) * sz, src + (g * n_data + s) * sz, sz);
		out = EGO->tmp;
	}
	char name[4096];
	snprintf(name, sizeof(name), o->outfile, (int)i);
	if (gray_write_raw(name, sz, n_data, o->w_rays, o->h_rays, out, EGO->info)) {
		fprintf(stderr, "GRay2: cannot write \"%s\"\n", name);
		return -1;
	}
	return 0;
}

/* sim/gray.c:153-185: n_dump intervals ending at dt_dump * (i + 1), each
 * followed by rayfile % (i + 1) holding /rays.  The rewrite leaves the
 * integration and the infcam initial conditions as TODOs
 * (sim/gray.c:168, sim/gray/infcam.c:40); here they are the pixel-centre
 * camera of sim-org/driver.cl:36-67 and the analytic Kerr integration. */
static int exec_rewrite(void *ego)
{
	struct gray_options *o = &EGO->o;
	const int w = (int)o->w_rays, h = (int)o->h_rays;
	double t = o->t_init;
	size_t i = o->i_init;
	if (grayhip_icond(EGO->ctx, o->w_img, o->h_img, o->r_obs, o->i_obs, o->j_obs)) {
		fprintf(stderr, "GRay2: icond: %s\n", grayhip_last_error(EGO->ctx));
		return EXIT_FAILURE;
	}
	while (i < o->n_dump) {
		const size_t next = i + 1;
		const double target = o->dt_dump * (double)next;
		double ns = 0.0;
		char name[4096];
		printf("%zu: %4.1f -> %4.1f", next, t, target);
		fflush(stdout);
		if (grayhip_evolve(EGO->ctx, target - t, o->n_sub, &ns) ||
		    grayhip_read_f64(EGO->ctx, EGO->data, NULL)) {
			fprintf(stderr, "\nGRay2: evolve: %s\n", grayhip_last_error(EGO->ctx));
			return EXIT_FAILURE;
		}
		snprintf(name, sizeof(name), o->rayfile, (int)next);
		if (gray_h5_write_rays(name, w, h, 8, EGO->data, (int)o->precision)) {
			fprintf(stderr, "\nGRay2: %s\n", gray_h5_error());
			return EXIT_FAILURE;
		}
		printf(": DONE\n");
		i = next;
		t = target;
	}
	return EXIT_SUCCESS;
}

static int _exec(void *ego)
{
	struct gray_options *o = &EGO->o;
	const size_t n_rays = o->w_rays * o->h_rays;
	size_t i = o->i_init;
	const double t_init = o->t_init, dt_dump = o->dt_dump;
	int rc;

	if (!EGO->ctx && !EGO->grid) {
		fprintf(stderr, "GRay2: exec before init\n");
		return EXIT_FAILURE;
	}
	if (rewrite(o))
		return exec_rewrite(ego);
	if (EGO->grid) {
		/* gray.c:164-241: open the file, list times, load the bracketing
		 * snapshots (or one, frozen) */
		if (gray_grid_load(EGO->grid, o->dyst_file, t_init, dt_dump, i, o->n_dump,
		                   o->enable_fast_light)) {
			fprintf(stderr, "GRay2: %s\n", gray_grid_error(EGO->grid));
			return EXIT_FAILURE;
		}
	}

	printf("%zu:  initialize at %4.1f", i, t_init);
	rc = EGO->ctx ? grayhip_icond(EGO->ctx, o->w_img, o->h_img, o->r_obs, o->i_obs, o->j_obs)
	              : gray_grid_icond(EGO->grid, o->w_img, o->h_img, o->r_obs, o->i_obs, o->j_obs);
	if (rc) {
		fprintf(stderr, "\nGRay2: icond: %s\n",
		        EGO->ctx ? grayhip_last_error(EGO->ctx) : gray_grid_error(EGO->grid));
		return EXIT_FAILURE;
	}
	if (dump(ego, i))
		return EXIT_FAILURE;
	printf(": DONE\n");

	while (i < o->n_dump) {
		double ns = 0.0, t, target;
		t = t_init + (double)i * dt_dump;
		target = t_init + (double)(++i) * dt_dump;
		printf("%zu: %4.1f -> %4.1f", i, t, target);
		fflush(stdout);
		rc = EGO->ctx ? grayhip_evolve(EGO->ctx, target - t, o->n_sub, &ns)
		              : gray_grid_evolve(EGO->grid, t, target, o->n_sub, &ns);
		if (rc) {
			fprintf(stderr, "\nGRay2: evolve: %s\n",
			        EGO->ctx ? grayhip_last_error(EGO->ctx) : gray_grid_error(EGO->grid));
			return EXIT_FAILURE;
		}
		if (dump(ego, i))
			return EXIT_FAILURE;
		printf(": DONE (%.3gns/step/ray)\n", ns / (double)o->n_sub / (double)n_rays);
	}
	return EXIT_SUCCESS;
}

void *gray_mkmod(const void *opts)
{
	(void)opts;
	struct gray *ego = calloc(1, sizeof(struct gray));
	if (ego) {
		ego->super.conf = _conf;
		ego->super.init = _init;
		ego->super.exec = _exec;
		gray_options_init(&ego->o);
		ego->n_coor = 8;
		ego->n_freq = 0;
		ego->n_info = 1;
	}
	return ego;
}

void gray_rmmod(void *ego)
{
	if (!ego)
		return;
	if (EGO->ctx)
		grayhip_destroy(EGO->ctx);
	if (EGO->grid)
		gray_grid_destroy(EGO->grid);
	free(EGO->data);
	free(EGO->info);
	free(EGO->tmp);
	gray_options_free(&EGO->o);
	free(ego);
}

int gray_conf(void *ego, const char *arg) { return ((gray_job_ops *)ego)->conf(ego, arg); }
int gray_init(void *ego) { return ((gray_job_ops *)ego)->init(ego); }
int gray_exec(void *ego) { return ((gray_job_ops *)ego)->exec(ego); }

size_t gray_n_freq(const void *ego) { return ((const struct gray *)ego)->n_freq; }

size_t gray_n_data(const void *ego)
{
	return ((const struct gray *)ego)->n_coor + 2 * ((const struct gray *)ego)->n_freq;
}

int gray_get_option(const void *ego, const char *name, char *buf, size_t n)
{
	return gray_options_get(&((const struct gray *)ego)->o, name, buf, n);
}

void gray_print_options(void *ego, void *stream)
{
	gray_options_print(&EGO->o, stream ? (FILE *)stream : stdout);
}

int gray_run(int argc, const char *const *argv)
{
	void *ego = gray_mkmod(NULL);
	int rc = EXIT_FAILURE;
	if (!ego)
		return EXIT_FAILURE;
	for (int k = 0; k < argc; ++k)
		if (gray_conf(ego, argv[k])) {
			fprintf(stderr, "GRay2: invalid option \"%s\"\n", argv[k]);
			goto done;
		}
	if (gray_init(ego) != EXIT_SUCCESS)
		goto done;
	rc = gray_exec(ego);
done:
	gray_rmmod(ego);
	return rc;
}
