/*
 * grid_oracle.c -- FP64 CPU restatement of GRay2's numerical-spacetime path
 * (coordinates=dyst).  TEST INFRASTRUCTURE ONLY (the checker).
 *
 * Follows sim-org/interp.cl:53-205 (cell location, linear-in-physical-
 * coordinate fraction, clamps, trilinear sampling, linear-in-time blend) and
 * sim-org/dyst.cl:55-194 (metric contraction, coordinate-time geodesic RHS),
 * with the decisions recorded for the undefined parts of the reference:
 *   - trilinear weights computed exactly in FP64 (the reference uses the
 *     OpenCL sampler's device-defined low-precision weights);
 *   - g_zz is read from g_zz (dyst.cl:78 reads g_yz);
 *   - the ray's grid time is t_off + q^t and the time weight is clamped to
 *     [0, 1] (interp.cl:204 extrapolates);
 *   - the initial direction is null-normalised with the interpolated metric
 *     (dyst.cl:89-110 does not normalise) unless raw_icond is set;
 *   - a ray stops when it is outside the grid box or its coordinate speed
 *     marks it as captured (the reference's dyst getrr/geteps stubs,
 *     dyst.cl:43-53, never stop it);
 *   - radiative transfer uses the reference flow (flow.cl:36-58) and
 *     emissivities (rt.cl, restated in rt_oracle.c).
 * Evaluation order is written out so the FP64 GPU kernel can match it.
 */
#include "grid_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* index of Gamma^mu_{ab} (a <= b) within the 10 per mu, file order
 * tt tx ty tz xx xy xz yy yz zz (sim-org/io.c:356-361) */
static const int sym_idx[4][4] = {
	{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9},
};

/* interp.cl:69-164 for one axis: cell i0 and fractional index u in [0,n-1]. */
static void locate(const og_grid *G, int ax, double x, int *i0, double *f)
{
	const int n = G->n[ax];
	const double lo = G->lo[ax], hi = G->hi[ax];
	double xi, x0, x1;
	int i;
	if (G->map == OG_MAP_FISHEYE) {
		/* x = sinh((A xi + B)^3): interp.cl:89-96, 115-120 */
		double B = cbrt(asinh(lo));
		double A = (cbrt(asinh(hi)) - B) / (double)(n - 1);
		xi = floor((cbrt(asinh(x)) - B) / A);
		i = xi >= 0.0 ? (xi <= (double)(n - 1) ? (int)xi : n - 1) : 0;
		double c0 = A * (double)i + B, c1 = A * (double)(i + 1) + B;
		x0 = sinh((c0 * c0) * c0);
		x1 = sinh((c1 * c1) * c1);
	} else {
		double dx = (hi - lo) / (double)(n - 1);
		xi = floor((x - lo) / dx);
		i = xi >= 0.0 ? (xi <= (double)(n - 1) ? (int)xi : n - 1) : 0;
		x0 = lo + (double)i * dx;
		x1 = lo + (double)(i + 1) * dx;
	}
	double u = (double)i + (x - x0) / (x1 - x0);
	if (!(u > 0.0))
		u = 0.0;                      /* also maps NaN to the edge */
	if (u > (double)(n - 1))
		u = (double)(n - 1);
	/* sampler: i0 = floor(u), fraction u - i0, upper index clamped */
	double fl = floor(u);
	*i0 = (int)fl;
	*f = u - fl;
}

/* Cell and trilinear weights of one point (shared by all fields). */
typedef struct {
	size_t c[8];     /* point indices of the 8 corners */
	double w[8];     /* their weights */
	double wt;       /* time weight, -1: space only */
} stencil;

static void make_stencil(const og_grid *G, const double q[4], stencil *S)
{
	int i[3];
	double f[3];
	for (int a = 0; a < 3; ++a)
		locate(G, a, q[a + 1], &i[a], &f[a]);
	int j[3][2];
	for (int a = 0; a < 3; ++a) {
		j[a][0] = i[a];
		j[a][1] = i[a] + 1 < G->n[a] ? i[a] + 1 : G->n[a] - 1;
	}
	double gx[2] = {1.0 - f[0], f[0]};
	double gy[2] = {1.0 - f[1], f[1]};
	double gz[2] = {1.0 - f[2], f[2]};
	const size_t nx = (size_t)G->n[0], nxy = nx * (size_t)G->n[1];
	int k = 0;
	for (int cz = 0; cz < 2; ++cz)
		for (int cy = 0; cy < 2; ++cy)
			for (int cx = 0; cx < 2; ++cx) {
				S->c[k] = (size_t)j[2][cz] * nxy + (size_t)j[1][cy] * nx + (size_t)j[0][cx];
				S->w[k] = (gx[cx] * gy[cy]) * gz[cz];
				++k;
			}
	if (G->t1 == G->t2) {
		S->wt = -1.0;
	} else {
		double w = ((G->t_off + q[0]) - G->t1) / (G->t2 - G->t1);
		S->wt = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
	}
}

/* One field (component comp of an interleaved array with stride ncomp) at
 * the stencil, blended in time (interp.cl:183-205). */
static double sample(const og_grid *G, const double *const lev[2], int ncomp, int comp,
                     const stencil *S)
{
	double y[2];
	int nl = S->wt < 0.0 ? 1 : 2;
	for (int l = 0; l < nl; ++l) {
		const double *a = lev[l];
		double s = S->w[0] * a[S->c[0] * ncomp + comp];
		for (int k = 1; k < 8; ++k)
			s = s + S->w[k] * a[S->c[k] * ncomp + comp];
		y[l] = s;
	}
	(void)G;
	if (nl == 1)
		return y[0];
	return y[0] + S->wt * (y[1] - y[0]);
}

/* dyst.cl:55-81 with g_zz fixed: k_mu = g_{mu nu} u^nu. */
void og_grid_down(const og_grid *G, const double q[4], const double u[4], double k[4])
{
	stencil S;
	double m[10], g[4][4];
	make_stencil(G, q, &S);
	for (int c = 0; c < 10; ++c)
		m[c] = sample(G, G->met, 10, c, &S);
	for (int a = 0; a < 4; ++a)
		for (int b = 0; b < 4; ++b)
			g[a][b] = m[sym_idx[a][b]];
	for (int a = 0; a < 4; ++a)
		k[a] = ((g[a][0] * u[0] + g[a][1] * u[1]) + g[a][2] * u[2]) + g[a][3] * u[3];
}

double og_grid_getuu(const og_grid *G, const double q[4], const double u[4])
{
	double k[4];
	og_grid_down(G, q, u, k);
	return ((u[0] * k[0] + u[1] * k[1]) + u[2] * k[2]) + u[3] * k[3];
}

/* dyst.cl:112-194: a^mu = -u.Gamma^mu.u + (u.Gamma^t.u) u^mu. */
static void rhs_at(const og_grid *G, const stencil *Sp, const double u[4], double du[4])
{
	const stencil S = *Sp;
	double gam[40], uGu[4];
	for (int c = 0; c < 40; ++c)
		gam[c] = sample(G, G->gam, 40, c, &S);
	for (int mu = 0; mu < 4; ++mu) {
		const double *Gm = gam + 10 * mu;
		double v[4];
		for (int a = 0; a < 4; ++a)
			v[a] = ((Gm[sym_idx[a][0]] * u[0] + Gm[sym_idx[a][1]] * u[1]) +
			        Gm[sym_idx[a][2]] * u[2]) + Gm[sym_idx[a][3]] * u[3];
		uGu[mu] = ((u[0] * v[0] + u[1] * v[1]) + u[2] * v[2]) + u[3] * v[3];
	}
	for (int mu = 0; mu < 4; ++mu)
		du[mu] = -uGu[mu] + uGu[0] * u[mu];
}

void og_grid_rhs(const og_grid *G, const double q[4], const double u[4], double du[4])
{
	stencil S;
	make_stencil(G, q, &S);
	rhs_at(G, &S, u, du);
}

/* flow.cl:36-58: n_e = rho interpolated like every other field; the fluid
 * is at rest, u_fluid = (1,0,0,0), so shift = -k.u_fluid = -k_t with
 * k_t = g_{t nu} u^nu (the t row of dyst.cl:55-81).  The field 4-vector is
 * b^mu = (0, 0, 0, B), B = b_field (flow.cl:47 sets b = 0, the default), so
 * k.b = k_z B with k_z from the z row of the metric, and bkcos =
 * k.b / (shift B + eps) (flow.cl:55); it is 0 for B = 0. */
static void flow_at(const og_grid *G, const stencil *S, const double u[4], double *ne,
                    double *shift, double *bkcos)
{
	double m[4];
	for (int c = 0; c < 4; ++c)
		m[c] = sample(G, G->met, 10, c, S);
	const double kt = ((m[0] * u[0] + m[1] * u[1]) + m[2] * u[2]) + m[3] * u[3];
	*shift = -kt;
	*ne = sample(G, G->rho, 1, 0, S);
	*bkcos = 0.0;
	if (G->b_field > 0.0) {
		const double gxz = sample(G, G->met, 10, 6, S);
		const double gyz = sample(G, G->met, 10, 8, S);
		const double gzz = sample(G, G->met, 10, 9, S);
		const double kz = ((m[3] * u[0] + gxz * u[1]) + gyz * u[2]) + gzz * u[3];
		*bkcos = (kz * G->b_field) / (*shift * G->b_field + OG_RT_EPS);
	}
}

void og_grid_flow(const og_grid *G, const double q[4], const double u[4], double *ne,
                  double *shift, double *bkcos)
{
	stencil S;
	make_stencil(G, q, &S);
	flow_at(G, &S, u, ne, shift, bkcos);
}

/* phys.cl:66-74 on a gridded spacetime: the geodesic part, then rt_rhs of
 * rt.cl:185-198 with the flow above (T_e = 1e12, B = b_field). */
void og_grid_rhs_all(const og_grid *G, const double *s, double *ds)
{
	stencil S;
	make_stencil(G, s, &S);
	ds[0] = s[4]; ds[1] = s[5]; ds[2] = s[6]; ds[3] = s[7];
	rhs_at(G, &S, s + 4, ds + 4);
	if (G->n_freq > 0) {
		double ne, shift, bkcos;
		flow_at(G, &S, s + 4, &ne, &shift, &bkcos);
		og_rt_thermal(G->rt, G->n_freq, G->nu, G->m_adm, ne, OG_RT_TE_FLOW, G->b_field, bkcos,
		              shift, s + 8, s + 8 + G->n_freq, ds + 8, ds + 8 + G->n_freq);
	}
}

/* dyst.cl:89-110 camera, plus (unless raw_icond) the null normalisation of
 * KS.cl:223-229 with the interpolated metric. */
void og_grid_icond(const og_grid *G, double r_obs, double si, double ci, double sj,
                   double cj, double alpha, double beta, double q[4], double u[4])
{
	double R0 = r_obs * si - beta * ci;
	double z  = r_obs * ci + beta * si;
	double y  = R0 * sj + alpha * cj;
	double x  = R0 * cj - alpha * sj;
	q[0] = 0.0; q[1] = x; q[2] = y; q[3] = z;
	u[0] = 1.0; u[1] = si * cj; u[2] = si * sj; u[3] = ci;
	if (G->raw_icond)
		return;
	stencil S;
	double m[10];
	make_stencil(G, q, &S);
	for (int c = 0; c < 10; ++c)
		m[c] = sample(G, G->met, 10, c, &S);
	/* m: tt tx ty tz xx xy xz yy yz zz */
	double A = m[0];
	double B = ((m[1] * u[1] + m[2] * u[2]) + m[3] * u[3]) * 2.0;
	double C = ((((m[4] * u[1] + m[5] * u[2]) + m[6] * u[3]) * u[1]) +
	            (((m[5] * u[1] + m[7] * u[2]) + m[8] * u[3]) * u[2])) +
	            (((m[6] * u[1] + m[8] * u[2]) + m[9] * u[3]) * u[3]);
	double s = (-(B + sqrt(B * B - (4.0 * A) * C))) / (2.0 * A);
	u[1] = u[1] / s;
	u[2] = u[2] / s;
	u[3] = u[3] / s;
}

/* phys.cl:48-64 with dyst's getrr = geteps = 1 (dyst.cl:43-53): the cap is
 * 0.01.  Unless ref_rule, the ray stops outside the grid box, and also where
 * its coordinate speed |dx/dt|^2 = u_x^2 + u_y^2 + u_z^2 (u^t = 1 in this
 * coordinate-time form) exceeds OG_GRID_U_CAP2: the horizon proxy.  Outside
 * a horizon light moves at |dx/dt| of order 1; deep inside, near the
 * singularity, the interpolated Christoffels make the speed grow without
 * bound until RK4 overflows. */
double og_grid_getdt(const og_grid *G, const double q[4], const double u[4], double dt)
{
	if (!G->ref_rule) {
		for (int a = 0; a < 3; ++a)
			if (!(q[a + 1] >= G->lo[a] && q[a + 1] <= G->hi[a]))
				return 0.0;
		const double v2 = (u[1] * u[1] + u[2] * u[2]) + u[3] * u[3];
		if (!(v2 <= OG_GRID_U_CAP2))
			return 0.0;
	}
	const double cap = (0.01 * 1.0) * 1.0;
	if (fabs(dt) > cap)
		return dt > 0.0 ? cap : -cap;
	return dt;
}

/* RK4.cl:37-47 on the whole state (geodesic and transfer components). */
void og_grid_integrate(const og_grid *G, double *s, double dt)
{
	const int nd = og_grid_n_data(G);
	double k1[24], k2[24], k3[24], k4[24], x[24];
	const double hdt = 0.5 * dt;
	og_grid_rhs_all(G, s, k1);
	for (int i = 0; i < nd; ++i) x[i] = s[i] + hdt * k1[i];
	og_grid_rhs_all(G, x, k2);
	for (int i = 0; i < nd; ++i) x[i] = s[i] + hdt * k2[i];
	og_grid_rhs_all(G, x, k3);
	for (int i = 0; i < nd; ++i) x[i] = s[i] + dt * k3[i];
	og_grid_rhs_all(G, x, k4);
	for (int i = 0; i < nd; ++i)
		s[i] = s[i] + (dt * ((k1[i] + 2.0 * (k2[i] + k3[i])) + k4[i])) / 6.0;
}

int64_t og_grid_evolve_ray(const og_grid *G, double *s, double dt, size_t n_sub, int *done)
{
	const int n = (int)((INT_MAX / (long)n_sub) * (long)n_sub);
	const double dts = dt / (double)n_sub;
	int64_t steps = 0;
	int h, dh;
	*done = 0;
	for (h = 0; h < n; h += dh) {
		double x = og_grid_getdt(G, s, s + 4, dts) / dt * (double)n;
		dh = (x == x) ? (int)x : 0;
		if (!dh) {
			*done = 1;
			break;
		}
		if (dh > n - h)
			dh = n - h;
		og_grid_integrate(G, s, (double)dh * dt / (double)n);
		++steps;
	}
	return steps;
}

void og_grid_icond_all(const og_grid *G, size_t w_rays, size_t h_rays, double w_img,
                       double h_img, double r_obs, double si, double ci, double sj, double cj,
                       double *data, double *info)
{
	const long n_rays = (long)(w_rays * h_rays);
#pragma omp parallel for schedule(static)
	for (long g = 0; g < n_rays; ++g) {
		size_t gi = (size_t)g % w_rays, gj = (size_t)g / w_rays;
		double alpha = (((double)gi + 0.5) / (double)w_rays - 0.5) * w_img;
		double beta  = (((double)gj + 0.5) / (double)h_rays - 0.5) * h_img;
		double *s = data + (size_t)g * (size_t)og_grid_n_data(G);
		og_grid_icond(G, r_obs, si, ci, sj, cj, alpha, beta, s, s + 4);
		for (int i = 8; i < og_grid_n_data(G); ++i)
			s[i] = 0.0;                   /* rt_icond, rt.cl:179-183 */
		if (info)
			info[g] = og_grid_getuu(G, s, s + 4);
	}
}

int64_t og_grid_evolve_all(const og_grid *G, size_t n_rays, double *data, double *info,
                           double dt, size_t n_sub, int64_t *steps_out, int n_threads)
{
	int64_t total = 0;
#ifdef _OPENMP
	if (n_threads > 0)
		omp_set_num_threads(n_threads);
#else
	(void)n_threads;
#endif
#pragma omp parallel for schedule(dynamic, 8) reduction(+:total)
	for (long g = 0; g < (long)n_rays; ++g) {
		double *s = data + (size_t)g * (size_t)og_grid_n_data(G);
		int done;
		int64_t st = og_grid_evolve_ray(G, s, dt, n_sub, &done);
		total += st;
		if (steps_out)
			steps_out[g] += st;
		if (info)
			info[g] = og_grid_getuu(G, s, s + 4);
	}
	return total;
}

double og_grid_sample(const og_grid *G, int which, int comp, const double q[4])
{
	stencil S;
	make_stencil(G, q, &S);
	if (which == 0)
		return sample(G, G->gam, 40, comp, &S);
	if (which == 2)
		return sample(G, G->rho, 1, 0, &S);
	return sample(G, G->met, 10, comp, &S);
}
