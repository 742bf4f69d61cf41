/*
 * gray_opts.h -- run-time options of the gray job.
 *
 * Names, aliases, types and defaults are those of the reference's option
 * tables sim-org/icond.opts, param.opts, setup.opts and sim/Kerr.opts (plus
 * the rewrite's sim/gray.opts and sim/infcam.opts names), which lux-optgen
 * turns into structs and key=value parsers.  A few options are added for
 * choices the reference leaves open (termination rule, fluid model); they
 * are marked "build" in the table.
 */
#ifndef GRAY_OPTS_H
#define GRAY_OPTS_H

#include <stddef.h>
#include <stdio.h>

enum { GRAY_DEV_DEFAULT = 1, GRAY_DEV_CPU = 2, GRAY_DEV_GPU = 4, GRAY_DEV_ACCEL = 8,
       GRAY_DEV_ALL = -1 };

struct gray_options {
	/* sim-org/icond.opts */
	double w_img, h_img, r_obs, i_obs, j_obs;
	/* sim-org/param.opts */
	double M_ADM;
	char  *coordinates;
	char  *dyst_file;
	size_t enable_fast_light;
	size_t w_rays, h_rays;
	double n_electron, beta_crit, R_high, R_low;
	double *nu;                 /* NaN-terminated list, NULL: RT off */
	/* sim-org/setup.opts */
	unsigned i_platform, i_device;
	int    device_type;
	size_t precision;
	char  *scheme, *morder, *kflags, *outfile;
	double t_init, dt_dump;
	size_t i_init, n_sub, n_dump;
	/* sim/Kerr.opts */
	double m_bh, a_spin;
	/* sim/gray.opts (rewrite) */
	char  *spacetime, *initcond, *rayfile;
	/* build additions */
	char  *escape;              /* "outward" (r<1.01 r_H or r>r_esc outward) | "ref" */
	double r_esc;
	double eps_stop;            /* <= 0: rule default */
	char  *fluid;               /* "disk" | "thermal" | "uniform" | "none"; "" by coordinates */
	double b_field;             /* thermal flow's b^z in Gauss (0: the reference flow) */
};

/* Fill defaults (lux-optgen's X_init). */
void gray_options_init(struct gray_options *o);
/* Free owned strings and lists. */
void gray_options_free(struct gray_options *o);
/* Parse one "key=value"; 0 when recognised and valid, non-zero otherwise.
 * On failure a reason is written to err (if not NULL). */
int  gray_options_conf(struct gray_options *o, const char *arg, char *err, size_t n);
/* Number of frequency channels in o->nu. */
size_t gray_options_n_freq(const struct gray_options *o);
/* Print the table with current values. */
void gray_options_print(const struct gray_options *o, FILE *f);
/* Current value of one option as text; 0 if found. */
int  gray_options_get(const struct gray_options *o, const char *name, char *buf, size_t n);

#endif
