/*
 * rt_thermal.h -- the reference's emission model on the device: thermal
 * bremsstrahlung with the piecewise Gaunt factor, thermal synchrotron
 * (Leung, Gammie & Noble 2011 eq. 72) and Kirchhoff absorption against the
 * Planck function (sim-org/rt.cl:80-170, 185-198), for the flow of
 * sim-org/flow.cl:36-58: n_e from the density, T_e = 1e12, fluid at rest,
 * field 4-vector b^mu = (0, 0, 0, B).  The reference flow has B = 0, for
 * which the synchrotron term is identically zero (its last factor is
 * n_e nu_s with nu_s proportional to B); it is then not evaluated.
 * bkcos = k.b / (shift B + eps) = k_z B / (shift B + eps) (flow.cl:55).
 *
 * T_e is a constant of the flow above the ln K_2 table (rt.cl:41-43), so
 * ln K_2 = ln(2 T_e^2 - 1/2) (rt.cl:165-167) is a host constant too.
 *
 * Constants that involve transcendental functions are computed once on the
 * host (rt_thermal_init), so both forms and the CPU oracle start from the
 * same values.
 *
 * FP64 form: the oracle's expressions in its evaluation order.
 * FP32 form: the emissivity carries the factor K = M_ADM ff_f^2
 * (~5e-32, with ff_f the constant of rt.cl:135) and the opacity K / B_nu;
 * both underflow FP32.  The FP32 form therefore integrates I' = I / K and
 * tau' = tau / K, with dI'/dl = -j' e^{-K tau'} / (shift^2 + eps) and
 * dtau'/dl = -j' shift / (B + eps), j' = n_e G / (sqrt(T_e) e^y + eps) * n_e,
 * and converts in FP64 when it loads and stores the state.
 */
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/grayhip.h"

namespace gray {

#define RT_EPS 1e-28          /* sim-org/preamble.cl:30 */
#define RT_TE_FLOW 1e12       /* sim-org/flow.cl:51     */

struct RtThermal {
	double planck1, planck2, c;                   /* 2hc, h/(m_e c), c          */
	double gaunt_a, gaunt_b, gaunt_c, gaunt_d, gaunt_e;
	double ff_x, ff_y, ff_f;                      /* m_e c^2/Ry, h/(m_e c^2), f */
	double m_adm, te, scale;                      /* M_ADM, T_e, K              */
	double nu[GRAYHIP_MAX_FREQ];
	int    n_freq;
	/* synchrotron (rt.cl:145-170); b = 0 turns it off */
	double b;                                     /* B, Gauss                   */
	double te2b;                                  /* (T_e T_e) B                */
	double syn_nus, syn_f, syn_k;                 /* e/(9 pi m_e c), f, 2^(11/12) */
	double e_lk2;                                 /* exp(-ln K_2) at T_e        */
};

/* Host: the constants of rt.cl in cgs (CODATA 2010), m_adm in solar masses,
 * b the flow's field in Gauss (0: the reference flow). */
void rt_thermal_init(RtThermal *r, double m_adm, double b, int n_freq, const double *nu);

/* ---------------- FP64: the oracle's sequence ---------------- */

__device__ __forceinline__ double b_planck_f64(const RtThermal &c, double nu, double te)
{
	const double n = nu / c.c;
	const double f1 = c.planck1 * (n * n);
	const double f2 = c.planck2 * (n / (te + RT_EPS));
	return n * (f2 > 1e-5 ? f1 / (exp(f2) - 1.0)
	                      : (f1 / f2) / ((1.0 + f2 / 2.0) + (f2 * f2) / 6.0));
}

__device__ __forceinline__ double gaunt_f64(const RtThermal &c, double x, double y)
{
	const double sx = sqrt(x), sy = sqrt(y);
	if (x > 1.0)
		return y > 1.0 ? c.gaunt_a / sy : c.gaunt_b * (c.gaunt_c - log(y + RT_EPS));
	if (x * y > 1.0)
		return c.gaunt_e / (sx * sy);
	if (y > sx)
		return 1.0;
	const double g = c.gaunt_b * (c.gaunt_d + log(sx / (y + RT_EPS)));
	return g > RT_EPS ? g : RT_EPS;
}

__device__ __forceinline__ double j_ff_f64(const RtThermal &c, double nu, double te, double ne)
{
	const double x = c.ff_x * te;
	const double y = c.ff_y * (nu / te);
	const double f = c.ff_f * ne;
	return ((c.m_adm * f) * gaunt_f64(c, x, y)) * (f / (sqrt(te) * exp(y) + RT_EPS));
}

/* rt.cl:145-170 at the flow's T_e (above T_MIN) and B = c.b. */
__device__ __forceinline__ double j_syn_f64(const RtThermal &c, double nu, double ne, double cs)
{
	if (cs <= -1.0 || cs >= 1.0)
		return 0.0;
	const double nus = (c.te2b * sqrt(1.0 - cs * cs)) * c.syn_nus;
	const double x = nu / (nus + RT_EPS);
	const double cx = cbrt(x);
	const double xx = sqrt(x) + c.syn_k * sqrt(cx);
	return ((c.m_adm * xx) * exp(-cx)) * (xx * c.e_lk2) * ((c.syn_f * ne) * nus);
}

/* rt.cl:185-198 for NF channels; bkcos as flow.cl:55. */
template <int NF>
__device__ __forceinline__ void thermal_f64(const RtThermal &c, double ne, double shift,
                                            double bkcos, const double *tau, double *dI,
                                            double *dtau)
{
#pragma unroll
	for (int i = 0; i < NF; ++i) {
		const double n = c.nu[i] * shift;
		const double B = b_planck_f64(c, n, c.te);
		const double L = c.b > 0.0 ? j_syn_f64(c, n, ne, bkcos) + j_ff_f64(c, n, c.te, ne)
		                           : j_ff_f64(c, n, c.te, ne);
		dI[i] = ((-L) * exp(-tau[i])) / (shift * shift + RT_EPS);
		dtau[i] = ((-L) * shift) / (B + RT_EPS);
	}
}

/* ---------------- FP32: scaled by K ---------------- */

struct RtThermalF {
	float planck1, planck2, inv_c;
	float gaunt_a, gaunt_b, gaunt_c, gaunt_d, gaunt_e;
	float ff_x_te, ff_y_over_te, sqrt_te, te_eps, scale;
	float nu[GRAYHIP_MAX_FREQ];
	/* synchrotron / K: nu_s = syn_nb sqrt(1 - cos^2), syn_nb = T_e^2 B e /
	 * (9 pi m_e c); j_syn / K = xx^2 e^{-cbrt x} syn_a n_e nu_s with
	 * syn_a = e^{-ln K_2} f_syn / ff_f^2 (M_ADM cancels against K) */
	float b, syn_nb, syn_k, syn_a;
	int   syn_on;
};

__host__ __device__ inline RtThermalF rt_thermal_f32(const RtThermal &c)
{
	RtThermalF f;
	f.planck1 = (float)c.planck1;
	f.planck2 = (float)c.planck2;
	f.inv_c = (float)(1.0 / c.c);
	f.gaunt_a = (float)c.gaunt_a;
	f.gaunt_b = (float)c.gaunt_b;
	f.gaunt_c = (float)c.gaunt_c;
	f.gaunt_d = (float)c.gaunt_d;
	f.gaunt_e = (float)c.gaunt_e;
	f.ff_x_te = (float)(c.ff_x * c.te);
	f.ff_y_over_te = (float)(c.ff_y / c.te);
	f.sqrt_te = (float)sqrt(c.te);
	f.te_eps = (float)(c.te + RT_EPS);
	f.scale = (float)c.scale;
	for (int i = 0; i < GRAYHIP_MAX_FREQ; ++i)
		f.nu[i] = (float)c.nu[i];
	f.syn_on = c.b > 0.0;
	f.b = (float)c.b;
	f.syn_nb = (float)(c.te2b * c.syn_nus);
	f.syn_k = (float)c.syn_k;
	f.syn_a = (float)(c.e_lk2 * (c.syn_f / (c.ff_f * c.ff_f)));
	return f;
}

/* j_syn / K in FP32 (rt.cl:145-170 with the constants above).  Where
 * cbrt(x) > 80, e^{-cbrt x} xx^2 < 1e-23 x and the term is dropped (x may
 * overflow FP32 there). */
__device__ __forceinline__ float j_syn_f32(const RtThermalF &c, float nu, float ne, float cs)
{
	if (cs <= -1.0f || cs >= 1.0f)
		return 0.0f;
	const float nus = c.syn_nb * __fsqrt_rn(1.0f - cs * cs);
	const float x = nu * __builtin_amdgcn_rcpf(nus + 1e-28f);
	const float cx = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f));
	if (!(cx < 80.0f))
		return 0.0f;
	const float xx = __fsqrt_rn(x) + c.syn_k * __fsqrt_rn(cx);
	return ((xx * __expf(-cx)) * xx) * c.syn_a * (ne * nus);
}

/* Channels i < nf of NF (nf uniform across the wave, as the serving pool
 * runs requests with fewer channels in a kernel built for more); the
 * others get zero derivatives. */
template <int NF>
__device__ __forceinline__ void thermal_f32(const RtThermalF &c, float ne, float shift,
                                            float bkcos, const float *tau_s, float *dI_s,
                                            float *dtau_s, int nf = NF)
{
	const float x = c.ff_x_te;
	const float inv_sh2 = 1.0f / (shift * shift + 1e-28f);
#pragma unroll
	for (int i = 0; i < NF; ++i) {
		if (i >= nf) {
			dI_s[i] = 0.f;
			dtau_s[i] = 0.f;
			continue;
		}
		const float nu = c.nu[i] * shift;
		/* B_nu at T_e (rt.cl:80-93) */
		const float n = nu * c.inv_c;
		const float f1 = c.planck1 * (n * n);
		const float f2 = c.planck2 * (n / c.te_eps);
		const float B = n * (f2 > 1e-5f ? f1 / (__expf(f2) - 1.0f)
		                                : (f1 / f2) / ((1.0f + 0.5f * f2) + (f2 * f2) * (1.0f / 6.0f)));
		/* Gaunt factor (rt.cl:95-120) */
		const float y = c.ff_y_over_te * nu;
		float g;
		if (x > 1.0f) {
			g = y > 1.0f ? c.gaunt_a * __frsqrt_rn(y) : c.gaunt_b * (c.gaunt_c - __logf(y + 1e-28f));
		} else {
			const float sx = __fsqrt_rn(x), sy = __fsqrt_rn(y);
			if (x * y > 1.0f)
				g = c.gaunt_e / (sx * sy);
			else if (y > sx)
				g = 1.0f;
			else {
				g = c.gaunt_b * (c.gaunt_d + __logf(sx / (y + 1e-28f)));
				g = g > 1e-28f ? g : 1e-28f;
			}
		}
		/* j_ff / K (rt.cl:122-143), plus j_syn / K with a field */
		float jp = (ne * g) * (ne / (c.sqrt_te * __expf(y) + 1e-28f));
		if (c.syn_on)
			jp = j_syn_f32(c, nu, ne, bkcos) + jp;
		dI_s[i] = -jp * __expf(-c.scale * tau_s[i]) * inv_sh2;
		dtau_s[i] = (-jp * shift) / (B + 1e-28f);
	}
}

} // namespace gray
