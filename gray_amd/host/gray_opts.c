/*
 * gray_opts.c -- table-driven key=value parser for the gray job options.
 *
 * One row per option: "name:alias..." (colon separated, as in the .opts
 * files), type, default text, and where it lives.  The table restates the
 * rows of sim-org/icond.opts:2-6, param.opts:4-18, setup.opts:2-16,
 * sim/Kerr.opts:1-2, sim/gray.opts:5-14 and sim/infcam.opts:9-11.
 */
#define _GNU_SOURCE
#include "gray_opts.h"

#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

typedef enum { T_REAL, T_SIZE, T_UINT, T_STR, T_REALS, T_PREC, T_DEVTYPE } opt_type;

typedef struct {
	const char *names;     /* "name:alias:alias" */
	opt_type    type;
	size_t      off;
	const char *def;       /* default, parsed by the same parser; NULL = none */
	const char *src;       /* where the row comes from */
	const char *doc;
} opt_row;

#define OFF(f) offsetof(struct gray_options, f)

static const opt_row rows[] = {
	/* sim-org/icond.opts:2-6 */
	{"w_img:w", T_REAL, OFF(w_img), "64", "icond.opts:2", "Width of the image in GM/c^2"},
	{"h_img:h", T_REAL, OFF(h_img), "64", "icond.opts:3", "Height of the image in GM/c^2"},
	{"r_obs:r", T_REAL, OFF(r_obs), "1024", "icond.opts:4", "Distance of the image from the black hole"},
	{"i_obs:i", T_REAL, OFF(i_obs), "60", "icond.opts:5", "Inclination angle of the image in degrees"},
	{"j_obs:j", T_REAL, OFF(j_obs), "0", "icond.opts:6", "Azimuthal angle of the image in degrees"},
	/* sim-org/param.opts:4-18 */
	{"M_ADM:m", T_REAL, OFF(M_ADM), "4.3e6", "param.opts:4", "Spacetime mass in solar masses"},
	{"coordinates:coor", T_STR, OFF(coordinates), "dyst", "param.opts:5",
	 "Coordinate system: KS (analytic Kerr-Schild) or dyst (numerical spacetime)"},
	{"dyst_file:data", T_STR, OFF(dyst_file), "data.h5", "param.opts:6",
	 "HDF5 file of the numerical spacetime"},
	{"enable_fast_light:fast_light", T_SIZE, OFF(enable_fast_light), "0", "param.opts:7",
	 "If non-zero, freeze the spacetime in time"},
	{"w_rays:W:n_width", T_SIZE, OFF(w_rays), "512", "param.opts:10", "Number of rays along the width"},
	{"h_rays:H:n_height", T_SIZE, OFF(h_rays), "512", "param.opts:11", "Number of rays along the height"},
	{"n_electron:ne", T_REAL, OFF(n_electron), "1e6", "param.opts:14", "Electron number density normalisation"},
	{"beta_crit:bc", T_REAL, OFF(beta_crit), "1", "param.opts:15", "Critical plasma beta"},
	{"R_high:Rh", T_REAL, OFF(R_high), "1", "param.opts:16", "T_p/T_e at high beta"},
	{"R_low:Rl", T_REAL, OFF(R_low), "1", "param.opts:17", "T_p/T_e at low beta"},
	{"nu", T_REALS, OFF(nu), NULL, "param.opts:18",
	 "Frequency channels (comma separated); empty turns radiative transfer off"},
	/* sim-org/setup.opts:2-16 */
	{"i_platform:ip", T_UINT, OFF(i_platform), "0", "setup.opts:2", "Index of platform (unused: HIP)"},
	{"i_device:id", T_UINT, OFF(i_device), "0", "setup.opts:3", "Index of the HIP device"},
	{"device_type:type", T_DEVTYPE, OFF(device_type), "ALL", "setup.opts:4", "Type of device (GPU is used)"},
	{"precision:p", T_PREC, OFF(precision), "4", "setup.opts:6",
	 "Size of real: 8 = plain FP64 form, 4 = FP32 form with FP64 fallback"},
	{"scheme:scm", T_STR, OFF(scheme), "RK4", "setup.opts:7", "Integration scheme (RK4)"},
	{"morder:mo:ml", T_STR, OFF(morder), "AoS", "setup.opts:8", "Memory order of the dump: AoS or SoA"},
	{"kflags:kf", T_STR, OFF(kflags), "-w", "setup.opts:9", "Kernel compilation flags (ignored: AOT build)"},
	{"outfile:out", T_STR, OFF(outfile), "%04d.raw", "setup.opts:11", "Dump file name format"},
	{"t_init:t0", T_REAL, OFF(t_init), "0", "setup.opts:12", "Initial time"},
	{"dt_dump:dt", T_REAL, OFF(dt_dump), "-32", "setup.opts:13", "Time interval between dumps"},
	{"i_init:i0", T_SIZE, OFF(i_init), "0", "setup.opts:14", "Initial dump id"},
	{"n_sub:n", T_SIZE, OFF(n_sub), "1024", "setup.opts:15", "Number of sub-steps per dump"},
	{"n_dump:N", T_SIZE, OFF(n_dump), "64", "setup.opts:16", "Number of dumps after the initial condition"},
	/* sim/Kerr.opts:1-2 */
	{"m_bh:M", T_REAL, OFF(m_bh), "4e6", "Kerr.opts:1", "Black-hole mass in solar masses"},
	{"a_spin:a", T_REAL, OFF(a_spin), "0", "Kerr.opts:2", "Dimensionless black-hole spin"},
	/* sim/gray.opts:5-14, sim/infcam.opts:9 (rewrite) */
	{"spacetime:st", T_STR, OFF(spacetime), "", "gray.opts:5",
	 "Kerr selects the rewrite's job: analytic Kerr, infcam rays, HDF5 ray files"},
	{"initcond:ic", T_STR, OFF(initcond), "infcam", "gray.opts:7", "Initial conditions (infcam)"},
	{"rayfile:ray", T_STR, OFF(rayfile), "ray%04d.h5", "gray.opts:14",
	 "Ray file name format of spacetime=Kerr (/rays, dims W x H x 8)"},
	/* build additions */
	{"escape:term", T_STR, OFF(escape), "outward", "build",
	 "Termination: outward (r<1.01 r_H, or r>r_esc moving outward) or ref (as phys.cl)"},
	{"r_esc", T_REAL, OFF(r_esc), "1000", "build", "Escape radius of the outward rule"},
	{"eps_stop", T_REAL, OFF(eps_stop), "0", "build", "Horizon stop r - r_H < eps_stop (0: rule default)"},
	{"fluid", T_STR, OFF(fluid), "", "build",
	 "Fluid for radiative transfer: disk (analytic default), thermal (the reference flow "
	 "on a numerical spacetime, its default), uniform or none"},
	{"b_field:bz", T_REAL, OFF(b_field), "0", "build",
	 "Field b^z of the thermal flow in Gauss (0: b = 0 as flow.cl; > 0 adds synchrotron)"},
};

#define N_ROWS (sizeof(rows) / sizeof(rows[0]))

static int name_matches(const char *names, const char *key, size_t klen)
{
	const char *p = names;
	while (*p) {
		const char *e = strchr(p, ':');
		size_t len = e ? (size_t)(e - p) : strlen(p);
		if (len == klen && strncmp(p, key, klen) == 0)
			return 1;
		if (!e)
			break;
		p = e + 1;
	}
	return 0;
}

static int parse_real(const char *v, double *out)
{
	char *rem;
	if (!*v)
		return -1;
	double x = strtod(v, &rem);
	while (isspace((unsigned char)*rem))
		++rem;
	if (*rem)
		return -1;
	*out = x;
	return 0;
}

static int parse_size(const char *v, size_t *out)
{
	char *rem;
	if (!*v || *v == '-')
		return -1;
	unsigned long long x = strtoull(v, &rem, 0);
	if (*rem)
		return -1;
	*out = (size_t)x;
	return 0;
}

/* NaN-terminated list of reals; "", "none", "null", "[]" give NULL. */
static int parse_reals(const char *v, double **out)
{
	size_t cap = 8, n = 0;
	const char *p = v;
	while (isspace((unsigned char)*p) || *p == '[')
		++p;
	if (!*p || *p == ']' || !strcasecmp(p, "none") || !strcasecmp(p, "null")) {
		*out = NULL;
		return 0;
	}
	double *a = malloc(cap * sizeof(double));
	if (!a)
		return -1;
	for (;;) {
		char *rem;
		while (isspace((unsigned char)*p))
			++p;
		double x = strtod(p, &rem);
		if (rem == p) {
			free(a);
			return -1;
		}
		if (n + 1 >= cap) {
			cap *= 2;
			double *b = realloc(a, cap * sizeof(double));
			if (!b) {
				free(a);
				return -1;
			}
			a = b;
		}
		a[n++] = x;
		p = rem;
		while (isspace((unsigned char)*p))
			++p;
		if (*p == ',' || *p == ';') {
			++p;
			continue;
		}
		if (*p == ']')
			++p;
		while (isspace((unsigned char)*p))
			++p;
		if (*p) {
			free(a);
			return -1;
		}
		break;
	}
	a[n] = NAN;
	*out = a;
	return 0;
}

static int parse_prec(const char *v, size_t *out)
{
	if (!strcmp(v, "4") || !strcasecmp(v, "single") || !strcasecmp(v, "float") ||
	    !strcasecmp(v, "f32") || !strcasecmp(v, "fp32"))
		*out = 4;
	else if (!strcmp(v, "8") || !strcasecmp(v, "double") || !strcasecmp(v, "f64") ||
	         !strcasecmp(v, "fp64"))
		*out = 8;
	else
		return -1;
	return 0;
}

static int parse_devtype(const char *v, int *out)
{
	static const struct { const char *s; int v; } t[] = {
		{"ALL", GRAY_DEV_ALL}, {"DEFAULT", GRAY_DEV_DEFAULT}, {"CPU", GRAY_DEV_CPU},
		{"GPU", GRAY_DEV_GPU}, {"ACCELERATOR", GRAY_DEV_ACCEL},
	};
	const char *s = v;
	if (!strncasecmp(s, "CL_DEVICE_TYPE_", 15))
		s += 15;
	for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i)
		if (!strcasecmp(s, t[i].s)) {
			*out = t[i].v;
			return 0;
		}
	return -1;
}

static int apply(struct gray_options *o, const opt_row *r, const char *v)
{
	char *base = (char *)o;
	switch (r->type) {
	case T_REAL:
		return parse_real(v, (double *)(base + r->off));
	case T_SIZE:
		return parse_size(v, (size_t *)(base + r->off));
	case T_UINT: {
		size_t x;
		if (parse_size(v, &x) || x > 0xffffffffu)
			return -1;
		*(unsigned *)(base + r->off) = (unsigned)x;
		return 0;
	}
	case T_STR: {
		char *s = strdup(v);
		if (!s)
			return -1;
		char **dst = (char **)(base + r->off);
		free(*dst);
		*dst = s;
		return 0;
	}
	case T_REALS: {
		double *a;
		if (parse_reals(v, &a))
			return -1;
		double **dst = (double **)(base + r->off);
		free(*dst);
		*dst = a;
		return 0;
	}
	case T_PREC:
		return parse_prec(v, (size_t *)(base + r->off));
	case T_DEVTYPE:
		return parse_devtype(v, (int *)(base + r->off));
	}
	return -1;
}

void gray_options_init(struct gray_options *o)
{
	memset(o, 0, sizeof(*o));
	for (size_t i = 0; i < N_ROWS; ++i)
		if (rows[i].def)
			(void)apply(o, &rows[i], rows[i].def);
}

void gray_options_free(struct gray_options *o)
{
	for (size_t i = 0; i < N_ROWS; ++i) {
		char *base = (char *)o;
		if (rows[i].type == T_STR) {
			free(*(char **)(base + rows[i].off));
			*(char **)(base + rows[i].off) = NULL;
		} else if (rows[i].type == T_REALS) {
			free(*(double **)(base + rows[i].off));
			*(double **)(base + rows[i].off) = NULL;
		}
	}
}

int gray_options_conf(struct gray_options *o, const char *arg, char *err, size_t n)
{
	const char *eq = arg ? strchr(arg, '=') : NULL;
	if (!eq) {
		if (err)
			snprintf(err, n, "\"%s\" is not of the form key=value", arg ? arg : "(null)");
		return 1;
	}
	size_t klen = (size_t)(eq - arg);
	for (size_t i = 0; i < N_ROWS; ++i) {
		if (!name_matches(rows[i].names, arg, klen))
			continue;
		if (apply(o, &rows[i], eq + 1)) {
			if (err)
				snprintf(err, n, "invalid value \"%s\" for %.*s", eq + 1, (int)klen, arg);
			return 2;
		}
		return 0;
	}
	if (err)
		snprintf(err, n, "unknown option \"%.*s\"", (int)klen, arg);
	return 1;
}

size_t gray_options_n_freq(const struct gray_options *o)
{
	size_t n = 0;
	if (o->nu)
		while (!isnan(o->nu[n]))
			++n;
	return n;
}

static void value_text(const struct gray_options *o, const opt_row *r, char *buf, size_t n)
{
	const char *base = (const char *)o;
	switch (r->type) {
	case T_REAL:
		snprintf(buf, n, "%.17g", *(const double *)(base + r->off));
		break;
	case T_SIZE:
	case T_PREC:
		snprintf(buf, n, "%zu", *(const size_t *)(base + r->off));
		break;
	case T_UINT:
		snprintf(buf, n, "%u", *(const unsigned *)(base + r->off));
		break;
	case T_DEVTYPE:
		snprintf(buf, n, "%d", *(const int *)(base + r->off));
		break;
	case T_STR: {
		const char *s = *(char *const *)(base + r->off);
		snprintf(buf, n, "%s", s ? s : "");
		break;
	}
	case T_REALS: {
		const double *a = *(double *const *)(base + r->off);
		size_t k = 0;
		buf[0] = '\0';
		for (size_t i = 0; a && !isnan(a[i]) && k + 1 < n; ++i)
			k += (size_t)snprintf(buf + k, n - k, i ? ",%.17g" : "%.17g", a[i]);
		break;
	}
	}
}

int gray_options_get(const struct gray_options *o, const char *name, char *buf, size_t n)
{
	for (size_t i = 0; i < N_ROWS; ++i)
		if (name_matches(rows[i].names, name, strlen(name))) {
			value_text(o, &rows[i], buf, n);
			return 0;
		}
	if (n)
		buf[0] = '\0';
	return -1;
}

void gray_options_print(const struct gray_options *o, FILE *f)
{
	char v[256];
	for (size_t i = 0; i < N_ROWS; ++i) {
		value_text(o, &rows[i], v, sizeof(v));
		fprintf(f, "  %-30s = %-14s %s [%s]\n", rows[i].names, v, rows[i].doc, rows[i].src);
	}
}
