/*
 * rt_oracle.c -- FP64 restatement of the transfer physics of
 * sim-org/rt.cl:28-198 and the flow of sim-org/flow.cl:36-58.
 * TEST INFRASTRUCTURE ONLY: used by the test suite, the smoke entry point
 * and the CPU-baseline leg of the benchmark, never by the product path.
 *
 * Formulas (and the published methods they follow):
 *   B_nu      Planck law 2 h nu^3 / c^2 / (e^x - 1), x = h nu / (k T), with
 *             x / (e^x - 1) ~ 1 / (1 + x/2 + x^2/6) for x <= 1e-5
 *             (rt.cl:80-93);
 *   j_ff      thermal bremsstrahlung, Rybicki & Lightman (1979) eq. 5.14b
 *             / 4 pi with Z = 1, n_i = n_e (rt.cl:122-143);
 *   gaunt     the piecewise velocity-averaged Gaunt factor of rt.cl:95-120
 *             (Rybicki & Lightman fig. 5.2 regions);
 *   j_syn     thermal synchrotron, Leung, Gammie & Noble (2011) eq. 72
 *             (rt.cl:145-170), with 1/K_2(1/theta) from a table of
 *             ln K_2 at 61 log-spaced theta in [0.1, 100] (rt.cl:51-78),
 *             regenerated here by quadrature, and ln(2 theta^2 - 1/2)
 *             above the table;
 *   transfer  dI/dl = -j e^-tau / (shift^2 + eps), dtau/dl = -j shift /
 *             (B + eps) per channel at nu = nu_i shift (rt.cl:185-198).
 * Built with -ffp-contract=off; each expression is parenthesised as it is
 * evaluated.
 */
#include "rt_oracle.h"

#include <math.h>

/* cgs physical constants (rt.cl:28-36; CODATA 2010) */
#define C_LIGHT 2.99792458e+10
#define H_PLANCK 6.62606957e-27
#define G_NEWTON 6.67384800e-08
#define K_BOLTZ 1.38064881e-16
#define RYDBERG 2.17987197e-11
#define Q_ELEC 4.80320425e-10
#define M_ELEC 9.10938291e-28
#define M_SUN 1.98910000e+33
#define ZETA_GAUNT 1.78107241799 /* e^gamma, gamma = 0.5772156649... */

double og_log_k2(double x)
{
	/* trapezoid on the whole line's half: the integrand is analytic and
	 * decays double-exponentially, so the rule converges geometrically in
	 * 1/h; h = 1/64 is far past double precision for x in [1e-3, 1e2]. */
	const double h = 1.0 / 64.0;
	/* scaled by e^x: K_2(x) e^x = int exp(-x (cosh t - 1)) cosh(2t) dt */
	double sum = 0.5;
	for (int k = 1; k < 1 << 20; ++k) {
		const double t = h * (double)k;
		const double e = x * (cosh(t) - 1.0);
		sum += exp(-e) * cosh(2.0 * t);
		if (e > 2.0 * t + 60.0)
			break;
	}
	return log(h * sum) - x;
}

void og_rt_consts_init(og_rt_consts *c)
{
	const double pi = 3.14159265358979323846;
	c->planck1 = (2.0 * H_PLANCK) * C_LIGHT;
	c->planck2 = H_PLANCK / (M_ELEC * C_LIGHT);
	c->c = C_LIGHT;
	c->gaunt_a = sqrt(3.0 / pi);
	c->gaunt_b = sqrt(3.0) / pi;
	c->gaunt_c = log(4.0 / ZETA_GAUNT);
	c->gaunt_d = log(4.0 / pow(ZETA_GAUNT, 2.5));
	c->gaunt_e = sqrt(12.0);
	c->ff_x = ((M_ELEC * C_LIGHT) * C_LIGHT) / RYDBERG;
	c->ff_y = H_PLANCK / ((M_ELEC * C_LIGHT) * C_LIGHT);
	c->ff_f = sqrt(((G_NEWTON * M_SUN) / (C_LIGHT * C_LIGHT)) * 6.8e-38 /
	               ((4.0 * pi) * sqrt(((M_ELEC * C_LIGHT) * C_LIGHT) / K_BOLTZ)));
	c->syn_nus = Q_ELEC / (((9.0 * pi) * M_ELEC) * C_LIGHT);
	c->syn_f = ((G_NEWTON * M_SUN) / (C_LIGHT * C_LIGHT)) *
	           ((sqrt(2.0) * pi) * (Q_ELEC * Q_ELEC) / (3.0 * C_LIGHT));
	c->syn_k = pow(2.0, 11.0 / 12.0);
	c->k2_h = (double)OG_RT_T_GRID / log(OG_RT_T_MAX / OG_RT_T_MIN);
	for (int i = 0; i <= OG_RT_T_GRID; ++i) {
		const double theta = OG_RT_T_MIN * pow(OG_RT_T_MAX / OG_RT_T_MIN,
		                                       (double)i / (double)OG_RT_T_GRID);
		c->log_k2[i] = og_log_k2(1.0 / theta);
	}
}

/* rt.cl:70-78; the cell index is clamped to the table (at te = T_MAX the
 * reference would read one entry past its end). */
double og_log_k2it(const og_rt_consts *c, double te)
{
	const double h = log(te / OG_RT_T_MIN) * c->k2_h;
	int i = (int)h;
	if (i > OG_RT_T_GRID - 1)
		i = OG_RT_T_GRID - 1;
	const double d = h - (double)i;
	return (1.0 - d) * c->log_k2[i] + d * c->log_k2[i + 1];
}

/* rt.cl:80-93 */
double og_b_planck(const og_rt_consts *c, double nu, double te)
{
	const double n = nu / c->c;
	const double f1 = c->planck1 * (n * n);
	const double f2 = c->planck2 * (n / (te + OG_RT_EPS));
	return n * (f2 > 1e-5 ? f1 / (exp(f2) - 1.0)
	                      : (f1 / f2) / ((1.0 + f2 / 2.0) + (f2 * f2) / 6.0));
}

/* rt.cl:95-120 */
double og_gaunt(const og_rt_consts *c, double x, double y)
{
	const double sx = sqrt(x), sy = sqrt(y);
	if (x > 1.0)
		return y > 1.0 ? c->gaunt_a / sy : c->gaunt_b * (c->gaunt_c - log(y + OG_RT_EPS));
	if (x * y > 1.0)
		return c->gaunt_e / (sx * sy);
	if (y > sx)
		return 1.0;
	const double g = c->gaunt_b * (c->gaunt_d + log(sx / (y + OG_RT_EPS)));
	return g > OG_RT_EPS ? g : OG_RT_EPS;
}

/* rt.cl:122-143 */
double og_j_ff(const og_rt_consts *c, double nu, double te, double ne, double m_adm)
{
	const double x = c->ff_x * te;
	const double y = c->ff_y * (nu / te);
	const double f = c->ff_f * ne;
	return ((m_adm * f) * og_gaunt(c, x, y)) * (f / (sqrt(te) * exp(y) + OG_RT_EPS));
}

/* rt.cl:145-170 */
double og_j_syn(const og_rt_consts *c, double nu, double te, double ne, double b,
                double cos_theta, double m_adm)
{
	if (te <= OG_RT_T_MIN || cos_theta <= -1.0 || cos_theta >= 1.0)
		return 0.0;
	const double nus = (((te * te) * b) * sqrt(1.0 - cos_theta * cos_theta)) * c->syn_nus;
	const double x = nu / (nus + OG_RT_EPS);
	const double cx = cbrt(x);
	const double xx = sqrt(x) + c->syn_k * sqrt(cx);
	const double lk2 = te > OG_RT_T_MAX ? log((2.0 * te) * te - 0.5) : og_log_k2it(c, te);
	return ((m_adm * xx) * exp(-cx)) * (xx * exp(-lk2)) * ((c->syn_f * ne) * nus);
}

/* rt.cl:185-198 */
void og_rt_thermal(const og_rt_consts *c, int n_freq, const double *nu, double m_adm,
                   double ne, double te, double b, double bkcos, double shift,
                   const double *I, const double *tau, double *dI, double *dtau)
{
	(void)I;
	for (int i = 0; i < n_freq; ++i) {
		const double n = nu[i] * shift;
		const double B = og_b_planck(c, n, te);
		const double L = og_j_syn(c, n, te, ne, b, bkcos, m_adm) + og_j_ff(c, n, te, ne, m_adm);
		dI[i] = ((-L) * exp(-tau[i])) / (shift * shift + OG_RT_EPS);
		dtau[i] = ((-L) * shift) / (B + OG_RT_EPS);
	}
}
