/*
 * step_f32.h -- one ray in the FP32 form: FP32 state derivative and the
 * classical RK4 step on a compensated (hi + lo) state.  P is the parameter
 * block (the render's RayParams or per-lane serving parameters).
 *
 * The synthetic disk's absorption is a fixed power of frequency, alpha_i
 * shift = 0.1 nu_i^-5/2 (n shift^-3/2), so every channel's optical depth is
 * tau_i = d_i T with d_i = 0.1 nu_i^-5/2 and one shared integral
 * T = -int n shift^-3/2 dlambda.  The FP32 form integrates T once instead of
 * the n_freq depths.
 *
 * Its geodesic is the Hamiltonian form (ks_ham): the state holds the
 * covariant momentum p_mu instead of u^mu, so a ray's state is t, x, y, z,
 * p_t, p_x, p_y, p_z, the n_freq intensities and T (ND32 reals); p_t is
 * conserved and not integrated.  load_state and store_state convert from
 * and to the stored layout (t, x, y, z, u^mu, intensities, tau_i in place)
 * in FP64 (down / up).
 */
#pragma once

#include "gray_device.h"

namespace gray {

/* Reals of the FP32 form's state: geodesic, intensities, T. */
template <int NF> constexpr int nd32() { return 8 + NF + (NF > 0 ? 1 : 0); }

/* Index of p_t in it (constant along the ray). */
constexpr int PT32 = 4;

/* A ray whose quantised step falls below this affine length (M) goes to the
 * FP64 re-integration.  Under the step rule's cap 0.01 (r - r_H)^2 that is
 * r - r_H < 0.022, the last stretch of a plunge before the stop at
 * r - r_H < 0.01: a step there still moves the position by tens of FP32
 * ulps, which the compensated state keeps.  With a threshold of 1e-4 (r -
 * r_H < 0.1) the rays winding at the prograde photon orbit of a = 0.9375
 * (r - r_H = 0.088) were re-integrated in FP64 too, and their chain of
 * ~1e5 sequential FP64 steps outlasted the pool. */
constexpr double FB_STEP_MIN = 5e-6;

/* Per-channel constants of the synthetic emissivity in base 2:
 * l2c_i = log2 nu_i^-1/2 and d2_i = 0.1 nu_i^-5/2 log2(e), so that a
 * channel's factor nu_i^-1/2 exp(-tau_i) = 2^(l2c_i - d2_i T) is one FMA
 * and one exp2. */
template <int NF> struct RtConst {
	float l2c[NF > 0 ? NF : 1], d2[NF > 0 ? NF : 1];
};

template <int NF, class P>
__host__ __device__ __forceinline__ RtConst<NF> make_rtconst(const P &p)
{
	RtConst<NF> r;
#pragma unroll
	for (int i = 0; i < (NF > 0 ? NF : 1); ++i) {
		const double nu = NF > 0 ? (double)p.nu[i] : 1.0;
		r.l2c[i] = (float)(-0.5 * log2(nu));
		r.d2[i] = (float)(0.1 / (nu * nu * sqrt(nu)) * 1.4426950408889634);
	}
	return r;
}

/* The transfer RHS of sim-org/rt.cl:185-198 with the synthetic disk,
 * evaluated from the state's covariant k_mu = p_mu and the Kerr-Schild
 * quantities the geodesic RHS already has: g(v,v) = eta(v,v) + f (l.v)^2
 * for the Keplerian v = (1, -y Omega, x Omega, 0); then
 *   dI_i = -n nu_i^-1/2 s^-1/2 exp(-tau_i) / (s^2 + 1e-28)
 *   dtau_i = -0.1 n nu_i^-5/2 s^-3/2       (s = -k.u_fluid),
 * with tau_i = d_i T and dT = -n s^-3/2; the exps run as exp2 of base-2
 * arguments (v_exp_f32, as __expf). */
template <int NF, class P>
__device__ __forceinline__ void rt_rhs_f32(const P &p, const RtConst<NF> &rc, const V4<float> &q,
                                           const V4<float> &k, const KsAux<float> &K, float T,
                                           float *dI, float &dT)
{
	typedef Mth<float> M;
	float n = 0.f, sh = 1.f;
	if (p.fluid == GRAYHIP_FLUID_DISK) {
		const float R2 = K.R2;
		if (K.r >= (float)p.r_isco && R2 > 0.f) {
			/* r^(3/2) and r^(-3/2) from one reciprocal square root */
			const float isr = __builtin_amdgcn_rsqf(K.r);
			const float rs = (K.r * K.r) * isr;
			const float Om = M::div(1.f, rs + (float)p.a);
			const float lv = 1.f + Om * (q.x * K.ly - q.y * K.lx);
			const float vv = (Om * Om * R2 - 1.f) + K.f * lv * lv;
			if (vv < 0.f) {
				const float ut = __builtin_amdgcn_rsqf(-vv);
				/* k.v = k_t + Omega (x k_y - y k_x), x k_y - y k_x = -B */
				const float kv = fmaf(-Om, K.B, k.t);
				const float s_ = -kv * ut;
				if (s_ > 0.f) {
					sh = s_;
					/* exp(-z^2 / (0.18 R^2)) as exp2 */
					n = (K.ir * isr) *
					    __builtin_amdgcn_exp2f(K.zz * __builtin_amdgcn_rcpf(R2) *
					                           (-1.4426950408889634f / 0.18f));
				}
			}
		}
	} else if (p.fluid == GRAYHIP_FLUID_UNIFORM) {
		n = 1.f;
	}
	if (n == 0.f) {
#pragma unroll
		for (int i = 0; i < NF; ++i)
			dI[i] = 0.f;
		dT = 0.f;
		return;
	}
	const float rsh = __builtin_amdgcn_rsqf(sh);
	const float A = n * rsh * M::div(1.f, sh * sh + 1e-28f);
	dT = -(n * (rsh * rsh) * rsh);
#pragma unroll
	for (int i = 0; i < NF; ++i)
		dI[i] = -A * __builtin_amdgcn_exp2f(fmaf(-rc.d2[i], T, rc.l2c[i]));
}

/* The FP32 state derivative; *r_out, when given, receives the Kerr-Schild
 * r of the state's position (for the step rule). */
template <int NF, class P>
__device__ __forceinline__ void rhs_f32(const P &p, const RtConst<NF> &rc, const float *s,
                                        float *ds, float *r_out = nullptr)
{
	V4<float> q = {s[0], s[1], s[2], s[3]};
	V4<float> k = {s[4], s[5], s[6], s[7]};
	KsAux<float> K;
	const Ham<float> h = ks_ham<float, HwF32>((float)p.a, q, k, NF > 0 ? &K : nullptr);
	ds[0] = h.u.t; ds[1] = h.u.x; ds[2] = h.u.y; ds[3] = h.u.z;
	ds[4] = 0.f;   ds[5] = h.dpx; ds[6] = h.dpy; ds[7] = h.dpz;
	if (r_out)
		*r_out = h.r;
	if constexpr (NF > 0)
		rt_rhs_f32<NF, P>(p, rc, q, k, K, s[8 + NF], ds + 8, ds[8 + NF]);
}

/* Classical RK4 (sim-org/RK4.cl:37-47) with stages taken from hi and the
 * increment added to (hi, lo) by compensated summation.  k1 is the first
 * stage, rhs_f32 at hi, which the caller has evaluated (its dx/dlambda gives
 * the step rule u^mu); p_t (PT32) is conserved and left as it is. */
template <int NF, class P>
__device__ __forceinline__ void rk4_f32c(const P &p, const RtConst<NF> &rc, float *hi, float *lo,
                                         const float *k1, float dt)
{
	constexpr int ND = nd32<NF>();
	float k[ND], acc[ND], x[ND];
	const float hdt = 0.5f * dt;

#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] = k1[i];
		x[i] = i == PT32 ? hi[i] : hi[i] + hdt * k1[i];
	}
	rhs_f32<NF, P>(p, rc, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		if (i == PT32)
			continue;
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + hdt * k[i];
	}
	rhs_f32<NF, P>(p, rc, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		if (i == PT32)
			continue;
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + dt * k[i];
	}
	rhs_f32<NF, P>(p, rc, x, k);
	const float dt6 = dt * (1.0f / 6.0f);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		if (i == PT32)
			continue;
		float y = (acc[i] + k[i]) * dt6 + lo[i];
		float t = hi[i] + y;
		lo[i] = y - (t - hi[i]);
		hi[i] = t;
	}
}

/* d_i = 0.1 nu_i^-5/2 in FP64 (tau_i = d_i T); channels a serving slot
 * carries beyond its request's repeat channel 0. */
template <class P>
__device__ __forceinline__ double tau_scale(const P &p, int i)
{
	const double nu = (double)p.nu[i < p.n_freq ? i : 0];
	return 0.1 / (nu * nu * sqrt(nu));
}

/* Stored layout (x: 8 + 2 NF reals) -> FP32-form state (hi + lo, ND32
 * reals): p_mu = down(u) at spin a, T from channel 0's depth. */
template <int NF, class P>
__device__ __forceinline__ void split_state(const P &p, double a, const double *x, float *hi,
                                            float *lo)
{
	const V4<double> q = {x[0], x[1], x[2], x[3]};
	const V4<double> k = down(a, q, V4<double>{x[4], x[5], x[6], x[7]});
	double y[8 + NF + 1];
	y[0] = x[0]; y[1] = x[1]; y[2] = x[2]; y[3] = x[3];
	y[4] = k.t;  y[5] = k.x;  y[6] = k.y;  y[7] = k.z;
#pragma unroll
	for (int i = 8; i < 8 + NF; ++i)
		y[i] = x[i];
	if constexpr (NF > 0)
		y[8 + NF] = x[8 + NF] / tau_scale(p, 0);
#pragma unroll
	for (int i = 0; i < nd32<NF>(); ++i) {
		hi[i] = (float)y[i];
		lo[i] = (float)(y[i] - (double)hi[i]);
	}
}

/* FP32-form state -> stored layout (8 + 2 NF reals: u^mu = up(p) at spin
 * a, tau_i = d_i T). */
template <int NF, class P>
__device__ __forceinline__ void join_state(const P &p, double a, const float *hi, const float *lo,
                                           double *x)
{
#pragma unroll
	for (int i = 0; i < 8 + NF; ++i)
		x[i] = (double)hi[i] + (double)lo[i];
	const V4<double> u = up(a, V4<double>{x[0], x[1], x[2], x[3]},
	                        V4<double>{x[4], x[5], x[6], x[7]});
	x[4] = u.t; x[5] = u.x; x[6] = u.y; x[7] = u.z;
	if constexpr (NF > 0) {
		const double T = (double)hi[8 + NF] + (double)lo[8 + NF];
#pragma unroll
		for (int i = 0; i < NF; ++i)
			x[8 + NF + i] = tau_scale(p, i) * T;
	}
}

template <int NF, class P>
__device__ __forceinline__ void load_state(const P &p, const double *__restrict__ data,
                                           size_t n_rays, size_t g, float *hi, float *lo)
{
	double x[8 + 2 * NF];
#pragma unroll
	for (int i = 0; i < 8 + NF + (NF > 0 ? 1 : 0); ++i)
		x[i] = data[(size_t)i * n_rays + g];
	split_state<NF, P>(p, p.a, x, hi, lo);
}

/* Writes the stored layout and returns g(u,u) of the stored state. */
template <int NF, class P>
__device__ __forceinline__ double store_state(const P &p, double *__restrict__ data,
                                              size_t n_rays, size_t g, const float *hi,
                                              const float *lo)
{
	double x[8 + 2 * NF];
	join_state<NF, P>(p, p.a, hi, lo, x);
#pragma unroll
	for (int i = 0; i < 8 + 2 * NF; ++i)
		data[(size_t)i * n_rays + g] = x[i];
	return getuu(p.a, V4<double>{x[0], x[1], x[2], x[3]}, V4<double>{x[4], x[5], x[6], x[7]});
}

} // namespace gray
