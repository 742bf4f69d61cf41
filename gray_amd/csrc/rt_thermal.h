/*
 * rt_thermal.h -- the reference's emission model on the device: thermal
 * bremsstrahlung with the piecewise Gaunt factor and Kirchhoff absorption
 * against the Planck function (sim-org/rt.cl:80-143, 185-198), for the flow
 * of sim-org/flow.cl:36-58 (n_e from the density, T_e = 1e12, b = 0, fluid
 * at rest).  The synchrotron term of rt.cl:145-170 is identically zero for
 * b = 0 (its last factor is n_e nu_s with nu_s proportional to b) and is
 * not evaluated.
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
};

/* Host: the constants of rt.cl in cgs (CODATA 2010), m_adm in solar masses. */
void rt_thermal_init(RtThermal *r, double m_adm, int n_freq, const double *nu);

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

/* rt.cl:185-198 for NF channels. */
template <int NF>
__device__ __forceinline__ void thermal_f64(const RtThermal &c, double ne, double shift,
                                            const double *tau, double *dI, double *dtau)
{
#pragma unroll
	for (int i = 0; i < NF; ++i) {
		const double n = c.nu[i] * shift;
		const double B = b_planck_f64(c, n, c.te);
		const double L = j_ff_f64(c, n, c.te, ne);
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
	return f;
}

/* Channels i < nf of NF (nf uniform across the wave, as the serving pool
 * runs requests with fewer channels in a kernel built for more); the
 * others get zero derivatives. */
template <int NF>
__device__ __forceinline__ void thermal_f32(const RtThermalF &c, float ne, float shift,
                                            const float *tau_s, float *dI_s, float *dtau_s,
                                            int nf = NF)
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
		/* j_ff / K (rt.cl:122-143) */
		const float jp = (ne * g) * (ne / (c.sqrt_te * __expf(y) + 1e-28f));
		dI_s[i] = -jp * __expf(-c.scale * tau_s[i]) * inv_sh2;
		dtau_s[i] = (-jp * shift) / (B + 1e-28f);
	}
}

} // namespace gray
