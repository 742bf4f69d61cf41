/*
 * gray_oracle.c -- FP64 CPU restatement of the GRay2 geodesic hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Built with
 * -ffp-contract=off so that every expression below is evaluated exactly
 * as parenthesised.  See gray_oracle.h for the conventions.
 *
 * The geodesic RHS and the RK4 stages fuse multiply-adds explicitly (FMA),
 * as an OpenCL compiler may under the default FP_CONTRACT ON (KS.cl counts
 * them), by one rule: in a sum taken left to right, a product added to the
 * running value is fused; where both operands are products, the left one is
 * rounded and the right one fused.  Built with -DOG_UNFUSED, FMA(a, b, c) is
 * a * b + c with two roundings (liboracle_unfused.so), to show that the
 * choice moves results far less than the parity tolerance.
 */
#include "gray_oracle.h"

#include <limits.h>
#include <math.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef OG_UNFUSED
#define FMA(a, b, c) ((a) * (b) + (c))
#else
#define FMA(a, b, c) fma((a), (b), (c))
#endif

/* OpenCL dot() restated left to right. */
static inline double dot3(double a0, double a1, double a2,
                          double b0, double b1, double b2)
{
	return (a0 * b0 + a1 * b1) + a2 * b2;
}

static inline double dot4(const double a[4], const double b[4])
{
	return ((a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]) + a[3] * b[3];
}

/* dot4 with every product after the first fused into the running sum. */
static inline double dot4f(const double a[4], const double b[4])
{
	return FMA(a[3], b[3], FMA(a[2], b[2], FMA(a[1], b[1], a[0] * b[0])));
}

void og_setup(og_params *p)
{
	/* sim-org/KS.cl:60: r_H = 1 + sqrt(1 - a^2) */
	p->r_h = 1.0 + sqrt(1.0 - p->a * p->a);
}

double og_r_isco(double a)
{
	/* Bardeen, Press & Teukolsky (1972) eq. 2.21, prograde orbits. */
	double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
	double z2 = sqrt(3.0 * a * a + z1 * z1);
	return 3.0 + z2 - sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

/* sim-org/KS.cl:48-55: r^2 from x^2+y^2+z^2 = r^2 + a^2 (1 - z^2/r^2). */
double og_getrr(double a, const double q[4])
{
	double aa = a * a;
	double zz = q[3] * q[3];
	double kk = 0.5 * (((q[1] * q[1] + q[2] * q[2]) + zz) - aa);
	return sqrt(kk * kk + aa * zz) + kk;
}

/* sim-org/KS.cl:57-61 */
double og_geteps(double a, const double q[4])
{
	return sqrt(og_getrr(a, q)) - (1.0 + sqrt(1.0 - a * a));
}

/* Kerr-Schild pieces shared by down() and gr_icond(): sim-org/KS.cl:94-108
 * and :207-221 are the same sequence. */
static void ks_metric(double a, const double q[4], double g[4][4])
{
	double aa = a * a;
	double zz = q[3] * q[3];
	double kk = 0.5 * (((q[1] * q[1] + q[2] * q[2]) + zz) - aa);
	double rr = sqrt(kk * kk + aa * zz) + kk;
	double r  = sqrt(rr);

	double f  = ((2.0 * rr) * r) / (rr * rr + aa * zz);
	double lx = (r * q[1] + a * q[2]) / (rr + aa);
	double ly = (r * q[2] - a * q[1]) / (rr + aa);
	double lz = q[3] / r;

	g[0][0] = -1.0 + f;  g[0][1] = f * lx;          g[0][2] = f * ly;          g[0][3] = f * lz;
	g[1][0] = f * lx;    g[1][1] = 1.0 + (f * lx) * lx; g[1][2] = (f * lx) * ly; g[1][3] = (f * lx) * lz;
	g[2][0] = f * ly;    g[2][1] = (f * ly) * lx;   g[2][2] = 1.0 + (f * ly) * ly; g[2][3] = (f * ly) * lz;
	g[3][0] = f * lz;    g[3][1] = (f * lz) * lx;   g[3][2] = (f * lz) * ly;   g[3][3] = 1.0 + (f * lz) * lz;
}

/* sim-org/KS.cl:91-114 */
void og_down(double a, const double q[4], const double u[4], double k[4])
{
	double g[4][4];
	ks_metric(a, q, g);
	k[0] = dot4(g[0], u);
	k[1] = dot4(g[1], u);
	k[2] = dot4(g[2], u);
	k[3] = dot4(g[3], u);
}

/* sim-org/KS.cl:164-168: dot(down(q,u), u) */
double og_getuu(double a, const double q[4], const double u[4])
{
	double k[4];
	og_down(a, q, u, k);
	return dot4(k, u);
}

/* sim-org/KS.cl:188-232.  The camera trigonometry (KS.cl:195-197) is an
 * input: si = sin(i), ci = cos(i), sj = sin(j), cj = cos(j). */
void og_gr_icond(double a, double r_obs, double si, double ci,
                 double sj, double cj, double alpha, double beta,
                 double q[4], double u[4])
{
	double R0 = r_obs * si - beta * ci;  /* cylindrical radius */
	double z  = r_obs * ci + beta * si;
	double y  = R0 * sj + alpha * cj;
	double x  = R0 * cj - alpha * sj;
	double g[4][4];

	q[0] = 0.0; q[1] = x; q[2] = y; q[3] = z;
	u[0] = 1.0; u[1] = si * cj; u[2] = si * sj; u[3] = ci;

	ks_metric(a, q, g);

	/* Null condition g_tt s^2 + 2 g_ti u^i s + g_ij u^i u^j = 0 solved
	 * for the spatial scale (KS.cl:223-229). */
	double A = g[0][0];
	double B = dot3(g[0][1], g[0][2], g[0][3], u[1], u[2], u[3]) * 2.0;
	double C = (dot3(g[1][1], g[1][2], g[1][3], u[1], u[2], u[3]) * u[1] +
	            dot3(g[2][1], g[2][2], g[2][3], u[1], u[2], u[3]) * u[2]) +
	            dot3(g[3][1], g[3][2], g[3][3], u[1], u[2], u[3]) * u[3];
	double s = (-(B + sqrt(B * B - (4.0 * A) * C))) / (2.0 * A);
	u[1] = u[1] / s;
	u[2] = u[2] / s;
	u[3] = u[3] / s;
}

/* sim-org/KS.cl:263-378: the regrouped Cartesian Kerr-Schild geodesic
 * equation  a^mu = -(eta^{mu b} u^a - eta^{mu a} u^b / 2) u_{b,a} + F l^mu. */
void og_gr_rhs(double a, const double q[4], const double u[4],
               double dq[4], double du[4])
{
	double f, dx_f, dy_f, dz_f;
	double lx, dx_lx, dy_lx, dz_lx;
	double ly, dx_ly, dy_ly, dz_ly;
	double lz, dx_lz, dy_lz, dz_lz;
	double tmp;

	/* f, l and their spatial gradients (KS.cl:278-335) */
	{
		double aa = a * a;
		double zz = q[3] * q[3];
		double kk = 0.5 * ((FMA(q[2], q[2], q[1] * q[1]) + zz) - aa);
		double dd = sqrt(FMA(aa, zz, kk * kk));
		double rr = dd + kk;
		double r  = sqrt(rr);
		double ir = 1.0 / r;
		double ss = rr + aa;
		double iss = 1.0 / ss;
		double dx_r, dy_r, dz_r, tmp2;

		tmp  = 0.5 / (r * dd);
		dz_r = (tmp * ss) * q[3];
		tmp  = tmp * rr;
		dy_r = tmp * q[2];
		dx_r = tmp * q[1];
		tmp  = 2.0 / (rr + (aa * zz) / rr);

		tmp2 = FMA(-(2.0 * rr), tmp, 3.0);
		f    = tmp * r;
		dx_f = (tmp * dx_r) * tmp2;
		dy_f = (tmp * dy_r) * tmp2;
		dz_f = tmp * FMA(-((tmp * aa) * q[3]), ir, dz_r * tmp2);

		double m2r  = -2.0 * r;
		double issr = iss * r;
		double issa = iss * a;

		lx    = iss * FMA(q[2], a, q[1] * r);
		tmp   = iss * FMA(m2r, lx, q[1]);
		dx_lx = FMA(tmp, dx_r, issr);
		dy_lx = FMA(tmp, dy_r, issa);
		dz_lx = tmp * dz_r;

		ly    = iss * FMA(-q[1], a, q[2] * r);
		tmp   = iss * FMA(m2r, ly, q[2]);
		dx_ly = FMA(tmp, dx_r, -issa);
		dy_ly = FMA(tmp, dy_r, issr);
		dz_ly = tmp * dz_r;

		lz    = q[3] * ir;
		tmp   = (-lz) * ir;
		dx_lz = tmp * dx_r;
		dy_lz = tmp * dy_r;
		dz_lz = FMA(tmp, dz_r, ir);
	}

	/* u_{b,a} contractions (KS.cl:337-366) */
	double hDxu, hDyu, hDzu, uD[4];
	{
		double Dx[4], Dy[4], Dz[4];
		double lu  = FMA(lz, u[3], FMA(ly, u[2], FMA(lx, u[1], u[0])));
		double flu = f * lu;
		Dx[0] = FMA(f, FMA(dx_lz, u[3], FMA(dx_ly, u[2], dx_lx * u[1])), dx_f * lu);
		Dy[0] = FMA(f, FMA(dy_lz, u[3], FMA(dy_ly, u[2], dy_lx * u[1])), dy_f * lu);
		Dz[0] = FMA(f, FMA(dz_lz, u[3], FMA(dz_ly, u[2], dz_lx * u[1])), dz_f * lu);

		Dx[1] = FMA(flu, dx_lx, Dx[0] * lx);
		Dx[2] = FMA(flu, dx_ly, Dx[0] * ly);
		Dx[3] = FMA(flu, dx_lz, Dx[0] * lz);

		Dy[1] = FMA(flu, dy_lx, Dy[0] * lx);
		Dy[2] = FMA(flu, dy_ly, Dy[0] * ly);
		Dy[3] = FMA(flu, dy_lz, Dy[0] * lz);

		Dz[1] = FMA(flu, dz_lx, Dz[0] * lx);
		Dz[2] = FMA(flu, dz_ly, Dz[0] * ly);
		Dz[3] = FMA(flu, dz_lz, Dz[0] * lz);

		hDxu = 0.5 * dot4f(Dx, u);
		hDyu = 0.5 * dot4f(Dy, u);
		hDzu = 0.5 * dot4f(Dz, u);

		for (int i = 0; i < 4; ++i)
			uD[i] = FMA(u[3], Dz[i], FMA(u[2], Dy[i], u[1] * Dx[i]));

		tmp = f * FMA(lz, uD[3] - hDzu, FMA(ly, uD[2] - hDyu, FMA(lx, uD[1] - hDxu, -uD[0])));
	}

	/* KS.cl:368-376 */
	dq[0] = u[0]; dq[1] = u[1]; dq[2] = u[2]; dq[3] = u[3];
	du[0] = uD[0] - tmp;
	du[1] = FMA(lx, tmp, hDxu - uD[1]);
	du[2] = FMA(ly, tmp, hDyu - uD[2]);
	du[3] = FMA(lz, tmp, hDzu - uD[3]);
}

/* sim-org/phys.cl:48-64.  r and eps are computed as the reference does
 * (getrr evaluated for r, then again inside geteps). */
double og_getdt(const og_params *p, const double q[4], const double u[4],
                double dt)
{
	double r   = sqrt(og_getrr(p->a, q));
	double eps = sqrt(og_getrr(p->a, q)) - p->r_h;

	if (eps < p->eps_stop)                   /* stop near horizon */
		return 0.0;

	if (p->escape == OG_ESC_OUTWARD) {       /* stop outside domain */
		double qu = dot3(q[1], q[2], q[3], u[1], u[2], u[3]);
		if (r > p->r_esc && qu * dt > 0.0)
			return 0.0;
	} else {
		/* phys.cl:57 tests dot(q,q) < 0, which never holds. */
		double qq = dot3(q[1], q[2], q[3], q[1], q[2], q[3]);
		if (qq < 0.0 && r > 1e3)
			return 0.0;
	}

	double cap = (0.01 * eps) * eps;
	if (fabs(dt) > cap)
		return dt > 0.0 ? cap : -cap;
	return dt;
}

/* Synthetic accretion disk: n = r^-1.5 exp(-z^2 /
 * (2 (0.3 rho)^2)), zero inside r_isco, on axis, or where the Keplerian
 * velocity is not timelike.  Keplerian Omega = 1/(r^1.5 + a) about the z
 * axis; in Cartesian Kerr-Schild coordinates v = (1, -y Omega, x Omega, 0)
 * and u_fluid = v / sqrt(-g(v,v)).  shift = -k_mu u_fluid^mu
 * (sim-org/flow.cl:54 with the static u replaced by the Keplerian one). */
void og_flow_disk(const og_params *p, const double q[4], const double k[4],
                  double *n, double *shift)
{
	double a  = p->a;
	double r  = sqrt(og_getrr(a, q));
	double R2 = q[1] * q[1] + q[2] * q[2];
	double rs = r * sqrt(r);

	*n = 0.0;
	*shift = 1.0;
	if (!(r >= p->r_isco) || !(R2 > 0.0))
		return;

	double Om = 1.0 / (rs + a);
	double v[4] = {1.0, -(q[2] * Om), q[1] * Om, 0.0};
	double vv = og_getuu(a, q, v);
	if (!(vv < 0.0))
		return;
	double ut = 1.0 / sqrt(-vv);
	double sh = -(dot4(k, v) * ut);
	if (!(sh > 0.0))
		return;

	*shift = sh;
	*n = (1.0 / rs) * exp(-(q[3] * q[3]) / (0.18 * R2));
}

/* sim-org/phys.cl:66-74 with rt_rhs (rt.cl:185-198):
 *   dI/dlambda   = -j_nu exp(-tau) / (shift^2 + 1e-28)
 *   dtau/dlambda = -alpha_nu shift
 * evaluated at the fluid-frame frequency nu = nu_i * shift, with the
 * synthetic emissivity j = n nu^-1/2 and absorption alpha = 0.1 n nu^-5/2. */
void og_rhs(const og_params *p, const double *s, double *ds)
{
	const int nf = p->n_freq;

	og_gr_rhs(p->a, s, s + 4, ds, ds + 4);
	if (nf == 0)
		return;

	double k[4], n = 0.0, shift = 1.0;
	og_down(p->a, s, s + 4, k);
	if (p->fluid == OG_FLUID_DISK)
		og_flow_disk(p, s, k, &n, &shift);
	else if (p->fluid == OG_FLUID_UNIFORM)
		n = 1.0;

	const double *tau = s + 8 + nf;
	double *dI = ds + 8, *dtau = ds + 8 + nf;
	if (n == 0.0) {
		for (int i = 0; i < nf; ++i) {
			dI[i] = 0.0;
			dtau[i] = 0.0;
		}
		return;
	}
	double den = shift * shift + 1e-28;
	for (int i = 0; i < nf; ++i) {
		double nu = p->nu[i] * shift;
		double sq = sqrt(nu);
		double j  = n / sq;
		double al = (0.1 * n) / ((nu * nu) * sq);
		dI[i]   = -(j * exp(-tau[i])) / den;
		dtau[i] = -(al * shift);
	}
}

/* sim-org/RK4.cl:37-47, elementwise over the whole state vector. */
void og_integrate(const og_params *p, double *s, double dt)
{
	const int nd = 8 + 2 * p->n_freq;
	double k1[8 + 2 * OG_MAX_FREQ], k2[8 + 2 * OG_MAX_FREQ];
	double k3[8 + 2 * OG_MAX_FREQ], k4[8 + 2 * OG_MAX_FREQ];
	double x[8 + 2 * OG_MAX_FREQ];
	double hdt = 0.5 * dt;

	og_rhs(p, s, k1);
	for (int i = 0; i < nd; ++i) x[i] = FMA(hdt, k1[i], s[i]);
	og_rhs(p, x, k2);
	for (int i = 0; i < nd; ++i) x[i] = FMA(hdt, k2[i], s[i]);
	og_rhs(p, x, k3);
	for (int i = 0; i < nd; ++i) x[i] = FMA(dt, k3[i], s[i]);
	og_rhs(p, x, k4);
	for (int i = 0; i < nd; ++i)
		s[i] = s[i] + (dt * (FMA(2.0, k2[i] + k3[i], k1[i]) + k4[i])) / 6.0;
}

/* sim-org/driver.cl:36-67 (pixel centre -> alpha, beta -> state). */
void og_icond_ray(const og_params *p, size_t w_rays, size_t h_rays,
                  double w_img, double h_img, double r_obs,
                  double si, double ci, double sj, double cj,
                  size_t gi, size_t gj, double *state)
{
	double alpha = (((double)gi + 0.5) / (double)w_rays - 0.5) * w_img;
	double beta  = (((double)gj + 0.5) / (double)h_rays - 0.5) * h_img;
	og_gr_icond(p->a, r_obs, si, ci, sj, cj, alpha, beta, state, state + 4);
	for (int i = 0; i < 2 * p->n_freq; ++i)
		state[8 + i] = 0.0;             /* rt_icond, rt.cl:179-183 */
}

/* sim-org/driver.cl:70-108 for one ray: integer-quantised sub-stepping.
 * n = (INT_MAX / n_sub) * n_sub units per launch; each step takes
 * dh = (int)(getdt / dt * n) units (0 retires the ray) clamped to the
 * units left, and advances the affine parameter by dh * dt / n. */
int64_t og_evolve_ray(const og_params *p, double *state, double dt,
                      size_t n_sub, int *done)
{
	const int n = (int)((INT_MAX / (long)n_sub) * (long)n_sub);
	const double dts = dt / (double)n_sub;
	int64_t steps = 0;
	int h, dh;

	*done = 0;
	for (h = 0; h < n; h += dh) {
		double x = og_getdt(p, state, state + 4, dts) / dt * (double)n;
		dh = (x == x) ? (int)x : 0;     /* NaN state retires the ray */
		if (!dh) {
			*done = 1;
			break;
		}
		if (dh > n - h)
			dh = n - h;
		og_integrate(p, state, (double)dh * dt / (double)n);
		++steps;
	}
	return steps;
}

void og_icond_all(const og_params *p, size_t w_rays, size_t h_rays,
                  double w_img, double h_img, double r_obs,
                  double si, double ci, double sj, double cj,
                  double *data, double *info)
{
	const size_t nd = 8 + 2 * (size_t)p->n_freq;
	const long n_rays = (long)(w_rays * h_rays);
#pragma omp parallel for schedule(static)
	for (long g = 0; g < n_rays; ++g) {
		size_t gi = (size_t)g % w_rays, gj = (size_t)g / w_rays;
		double *s = data + (size_t)g * nd;
		og_icond_ray(p, w_rays, h_rays, w_img, h_img, r_obs,
		             si, ci, sj, cj, gi, gj, s);
		if (info)
			info[g] = og_getuu(p->a, s, s + 4);
	}
}

int64_t og_evolve_all(const og_params *p, size_t n_rays, double *data,
                      double *info, double dt, size_t n_sub,
                      int64_t *steps_out, int n_threads)
{
	const size_t nd = 8 + 2 * (size_t)p->n_freq;
	int64_t total = 0;
#ifdef _OPENMP
	if (n_threads > 0)
		omp_set_num_threads(n_threads);
#else
	(void)n_threads;
#endif
#pragma omp parallel for schedule(dynamic, 16) reduction(+:total)
	for (long g = 0; g < (long)n_rays; ++g) {
		double *s = data + (size_t)g * nd;
		int done;
		int64_t st = og_evolve_ray(p, s, dt, n_sub, &done);
		total += st;
		if (steps_out)
			steps_out[g] += st;
		if (info)
			info[g] = og_getuu(p->a, s, s + 4);
	}
	return total;
}

int64_t og_evolve_subset(const og_params *p, size_t n_rays, size_t stride,
                         double *data, double dt, size_t n_sub,
                         size_t n_dump, int n_threads)
{
	const size_t nd = 8 + 2 * (size_t)p->n_freq;
	const long m = (long)((n_rays + stride - 1) / stride);
	int64_t total = 0;
#ifdef _OPENMP
	if (n_threads > 0)
		omp_set_num_threads(n_threads);
#else
	(void)n_threads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+:total)
	for (long j = 0; j < m; ++j) {
		double *s = data + (size_t)j * stride * nd;
		for (size_t d = 0; d < n_dump; ++d) {
			int done;
			total += og_evolve_ray(p, s, dt, n_sub, &done);
			if (done)
				break;
		}
	}
	return total;
}

/* KS.cl:63-71 in the spheroidal convention of the metric of
 * KS.cl:100-108, x + i y = (r + i a) e^{i phi} sin(theta): phi =
 * atan2(y r - x a, x r + y a).  (KS.cl:68-69 inverts the mirrored
 * convention, r - i a, in which this metric has no r-only map to
 * Boyer-Lindquist coordinates.) */
void og_getsphks(double a, const double q[4], double out[4])
{
	const double r = sqrt(og_getrr(a, q));
	out[0] = q[0];
	out[1] = r;
	out[2] = acos(q[3] / r);
	out[3] = atan2(q[2] * r - q[1] * a, q[1] * r + q[2] * a);
}

/* KS.cl:73-89 */
void og_getbl(double a, double r_match, const double q[4], double out[4])
{
	double s[4];
	og_getsphks(a, q, s);
	const double h = sqrt(1.0 - a * a), rp = 1.0 + h, rm = 1.0 - h, r = s[1];
	out[0] = s[0] + (rm * log((r - rm) / (r_match - rm)) - rp * log((r - rp) / (r_match - rp))) / h;
	out[1] = s[1];
	out[2] = s[2];
	out[3] = s[3] + ((0.5 * a) * log((r - rm) / (r - rp))) / h;
}

/* KS.cl:113-131: J1 = (0, rr x, rr y, (rr + aa) z) / (2 r dd) = dr,
 * J2 = ((z / r) J1 - (0,0,0,1)) / sqrt(rr - zz) = dtheta,
 * J3 = (a / (rr + aa)) J1 + (0, -y, x, 0) / RR = dphi of og_getsphks
 * (KS.cl:125 has -a, the derivative of its mirrored phi). */
void og_getsphksu(double a, const double q[4], const double u[4], double out[4])
{
	const double aa = a * a;
	const double RR = q[1] * q[1] + q[2] * q[2];
	const double zz = q[3] * q[3];
	const double kk = 0.5 * ((RR + zz) - aa);
	const double dd = sqrt(kk * kk + aa * zz);
	const double rr = dd + kk;
	const double r = sqrt(rr);
	const double d = (2.0 * r) * dd;
	const double J1[4] = {0.0, (rr * q[1]) / d, (rr * q[2]) / d, ((rr + aa) * q[3]) / d};
	const double sz = sqrt(rr - zz);
	const double J2[4] = {((q[3] / r) * J1[0]) / sz, ((q[3] / r) * J1[1]) / sz,
	                      ((q[3] / r) * J1[2]) / sz, ((q[3] / r) * J1[3] - 1.0) / sz};
	const double c = a / (rr + aa);
	const double J3[4] = {c * J1[0], c * J1[1] + (-q[2]) / RR, c * J1[2] + q[1] / RR, c * J1[3]};
	out[0] = u[0];
	out[1] = ((J1[0] * u[0] + J1[1] * u[1]) + J1[2] * u[2]) + J1[3] * u[3];
	out[2] = ((J2[0] * u[0] + J2[1] * u[1]) + J2[2] * u[2]) + J2[3] * u[3];
	out[3] = ((J3[0] * u[0] + J3[1] * u[1]) + J3[2] * u[2]) + J3[3] * u[3];
}

/* KS.cl:133-149 with u^r in the t and phi corrections (see the header). */
void og_getblu(double a, const double q[4], const double u[4], double out[4])
{
	double s[4];
	og_getsphksu(a, q, u, s);
	const double aa = a * a;
	const double zz = q[3] * q[3];
	const double kk = 0.5 * (((q[1] * q[1] + q[2] * q[2]) + zz) - aa);
	const double rr = sqrt(kk * kk + aa * zz) + kk;
	const double r = sqrt(rr);
	const double D = (rr - 2.0 * r) + aa;
	out[0] = s[0] - (s[1] * (2.0 * r)) / D;
	out[1] = s[1];
	out[2] = s[2];
	out[3] = s[3] - (s[1] * a) / D;
}
