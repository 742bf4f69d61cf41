/*
 * gray_bks.c -- boosted Kerr-Schild spacetime sampled on a grid: the input
 * generator of the numerical-spacetime path (the role of the reference's
 * tools/generate_data.py with tools/boosted_ks.py).
 *
 * Metric (tools/boosted_ks.py:120-186, restated): in the rest frame of the
 * hole, g = eta + 2H l l with H = M r^3 / (r^4 + a^2 z0^2) and the null
 * covector l = (1, (r x0 + a y0)/(r^2+a^2), (r y0 - a x0)/(r^2+a^2), z0/r).
 * The hole moves with speed v along the unit axis n (the reference boosts
 * along z only); rest-frame coordinates are x0 = Lambda x with
 *   t0 = gamma (t - v n.x),  x0 = x + (gamma - 1)(n.x) n - gamma v t n,
 * and the lab covector is l_lab = Lambda^T l0, so g_lab = eta + 2H l_lab l_lab.
 * Derivatives of g come from forward-mode dual numbers carrying all four
 * partials at once (the reference differentiates with one-slot duals,
 * boosted_ks.py:40-115,191-228); the inverse is g^-1 = eta - 2H l^ l^ with
 * l^ = eta l_lab (Kerr-Schild form), and
 *   Gamma^a_{mn} = 1/2 g^{ab} (d_m g_bn + d_n g_bm - d_b g_mn).
 * Non-finite samples become 0 (NaN) or +-DBL_MAX (inf), like the
 * reference's numpy.nan_to_num (generate_data.py:150-158).
 */
#include "../../include/gray_spacetime.h"

#include <float.h>
#include <math.h>
#include <string.h>

typedef struct { double v, d[4]; } dual;

static inline dual dc(double c) { dual r = {c, {0, 0, 0, 0}}; return r; }
static inline dual dadd(dual a, dual b) { dual r; r.v = a.v + b.v; for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
static inline dual dsub(dual a, dual b) { dual r; r.v = a.v - b.v; for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
static inline dual dmul(dual a, dual b) { dual r; r.v = a.v * b.v; for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
static inline dual dscale(dual a, double s) { dual r; r.v = a.v * s; for (int i = 0; i < 4; ++i) r.d[i] = a.d[i] * s; return r; }
static inline dual ddiv(dual a, dual b)
{
	dual r;
	r.v = a.v / b.v;
	for (int i = 0; i < 4; ++i)
		r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
	return r;
}
static inline dual dsqrt(dual a)
{
	dual r;
	r.v = sqrt(a.v);
	for (int i = 0; i < 4; ++i)
		r.d[i] = a.d[i] / (2.0 * r.v);
	return r;
}

/* Covector l_lab and H at lab event x (duals carry d/dx^mu). */
static void bks_l(const gray_bks *b, const dual x[4], dual l[4], dual *H)
{
	const double v = b->v, gam = 1.0 / sqrt(1.0 - v * v);
	const double *n = b->n;
	dual nx = dadd(dadd(dscale(x[1], n[0]), dscale(x[2], n[1])), dscale(x[3], n[2]));
	/* rest-frame coordinates */
	dual x0[3];
	for (int i = 0; i < 3; ++i)
		x0[i] = dsub(dadd(x[i + 1], dscale(nx, (gam - 1.0) * n[i])), dscale(x[0], gam * v * n[i]));
	const double a = b->a, aa = a * a;
	dual rho2 = dadd(dadd(dmul(x0[0], x0[0]), dmul(x0[1], x0[1])), dmul(x0[2], x0[2]));
	dual k = dscale(dsub(rho2, dc(aa)), 0.5);
	dual r2 = dadd(k, dsqrt(dadd(dmul(k, k), dscale(dmul(x0[2], x0[2]), aa))));
	dual r = dsqrt(r2);
	dual s = dadd(r2, dc(aa));
	*H = ddiv(dscale(dmul(r2, r), b->m), dadd(dmul(r2, r2), dscale(dmul(x0[2], x0[2]), aa)));
	dual l0t = dc(1.0);
	dual l0[3];
	l0[0] = ddiv(dadd(dmul(r, x0[0]), dscale(x0[1], a)), s);
	l0[1] = ddiv(dsub(dmul(r, x0[1]), dscale(x0[0], a)), s);
	l0[2] = ddiv(x0[2], r);
	dual nl = dadd(dadd(dscale(l0[0], n[0]), dscale(l0[1], n[1])), dscale(l0[2], n[2]));
	/* l_lab = Lambda^T l0 */
	l[0] = dsub(dscale(l0t, gam), dscale(nl, gam * v));
	for (int j = 0; j < 3; ++j)
		l[j + 1] = dadd(dadd(dscale(l0t, -gam * v * n[j]), l0[j]), dscale(nl, (gam - 1.0) * n[j]));
}

static const int sym10[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
static const double eta[4] = {-1.0, 1.0, 1.0, 1.0};

void gray_bks_setup(gray_bks *b, double m, double a, double v, const double axis[3])
{
	double nn = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
	b->m = m;
	b->a = a;
	b->v = v;
	for (int i = 0; i < 3; ++i)
		b->n[i] = nn > 0.0 ? axis[i] / nn : (i == 2 ? 1.0 : 0.0);
}

void gray_bks_point(const gray_bks *b, const double xe[4], double met[10], double gam[40])
{
	dual x[4], l[4], H;
	for (int mu = 0; mu < 4; ++mu) {
		x[mu] = dc(xe[mu]);
		x[mu].d[mu] = 1.0;
	}
	bks_l(b, x, l, &H);
	/* g_mn = eta_mn + 2 H l_m l_n with its four partials */
	dual g[10];
	for (int m = 0; m < 4; ++m)
		for (int n = m; n < 4; ++n) {
			dual t = dscale(dmul(H, dmul(l[m], l[n])), 2.0);
			if (m == n)
				t.v += eta[m];
			g[sym10[m][n]] = t;
		}
	for (int c = 0; c < 10; ++c)
		met[c] = g[c].v;
	if (!gam)
		return;
	/* g^{ab} = eta^{ab} - 2H l^a l^b */
	double lu[4], gi[4][4];
	for (int a = 0; a < 4; ++a)
		lu[a] = eta[a] * l[a].v;
	for (int a = 0; a < 4; ++a)
		for (int bb = 0; bb < 4; ++bb)
			gi[a][bb] = (a == bb ? eta[a] : 0.0) - 2.0 * H.v * lu[a] * lu[bb];
	/* first-kind symbols [b; m n] = 1/2 (d_m g_bn + d_n g_bm - d_b g_mn) */
	for (int a = 0; a < 4; ++a)
		for (int m = 0; m < 4; ++m)
			for (int n = m; n < 4; ++n) {
				double s = 0.0;
				for (int bb = 0; bb < 4; ++bb) {
					double first = 0.5 * (g[sym10[bb][n]].d[m] + g[sym10[bb][m]].d[n] -
					                       g[sym10[m][n]].d[bb]);
					s += gi[a][bb] * first;
				}
				gam[10 * a + sym10[m][n]] = s;
			}
}

static double clean(double x)
{
	if (isnan(x))
		return 0.0;
	if (isinf(x))
		return x > 0 ? DBL_MAX : -DBL_MAX;
	return x;
}

void gray_bks_sample(const gray_bks *b, double t, int nx, int ny, int nz, const double *xs,
                     const double *ys, const double *zs, double *gam, double *met, double *rho)
{
	const long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
	for (long p = 0; p < n; ++p) {
		int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((long)nx * ny));
		double xe[4] = {t, xs[i], ys[j], zs[k]};
		double mloc[10], gloc[40];
		gray_bks_point(b, xe, mloc, gloc);
		if (met)
			for (int c = 0; c < 10; ++c)
				met[p * 10 + c] = clean(mloc[c]);
		if (gam)
			for (int c = 0; c < 40; ++c)
				gam[p * 40 + c] = clean(gloc[c]);
		if (rho)
			rho[p] = clean(1.0 / sqrt(xe[1] * xe[1] + xe[2] * xe[2] + xe[3] * xe[3]));
	}
}

void gray_axis_uniform(double lo, double hi, int n, double *out)
{
	const double dx = (hi - lo) / (double)(n - 1);
	for (int i = 0; i < n; ++i)
		out[i] = lo + (double)i * dx;
	out[n - 1] = hi;
}

void gray_axis_fisheye(double lo, double hi, int n, double *out)
{
	/* x = sinh((A xi + B)^3), tools/generate_data.py:49-69 / interp.cl:115-120 */
	const double B = cbrt(asinh(lo));
	const double A = (cbrt(asinh(hi)) - B) / (double)(n - 1);
	for (int i = 0; i < n; ++i) {
		double c = A * (double)i + B;
		out[i] = sinh((c * c) * c);
	}
}
