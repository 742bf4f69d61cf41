/*
 * step_fb.h -- the FP64 step of the FP32 form's re-integration (rays whose
 * FP32 step collapses, handed over by a pool or a serving epoch).
 *
 * These rays are not part of the bit-exact form, and their dependency chain
 * of sequential steps is the critical path of a render's tail, so the step
 * is built for few instructions per step at FP64 accuracy:
 *   - divisions and square roots from the hardware FP64 reciprocal and
 *     reciprocal square root refined by two Newton steps (about 1 ulp
 *     instead of correctly rounded);
 *   - the transfer RHS in the regrouped form of the FP32 form: one
 *     reciprocal square root of the redshift and one exp per channel, with
 *     the per-channel constants nu_i^-1/2 and 0.1 nu_i^-5/2 formed once;
 *     k.u_fluid from the Kerr-Schild quantities gr_rhs already has;
 *   - the RK4 combination multiplied by 1/6.
 * The arithmetic is the same mathematics as sim-org/KS.cl:263-378,
 * rt.cl:185-198 and RK4.cl:37-47 in a different evaluation order.
 */
#pragma once

#include "gray_device.h"

namespace gray {

/* ~1 ulp FP64 division and square root. */
struct MthFb {
	static __device__ __forceinline__ double rcp(double b)
	{
		double r = __builtin_amdgcn_rcp(b);
		r = __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
		return __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
	}
	static __device__ __forceinline__ double rsq(double x)
	{
		double y = __builtin_amdgcn_rsq(x);
		/* y (3 - x y^2) / 2, twice */
		y = y * __builtin_fma(-0.5 * x * y, y, 1.5);
		return y * __builtin_fma(-0.5 * x * y, y, 1.5);
	}
	static __device__ __forceinline__ double div(double a, double b) { return a * rcp(b); }
	static __device__ __forceinline__ double sqrt_(double x)
	{
		return x > 0.0 ? x * rsq(x) : __builtin_sqrt(x);
	}
	static __device__ __forceinline__ double exp_(double x) { return exp(x); }
};

template <int NF> struct RtConstFb {
	double c[NF > 0 ? NF : 1], d[NF > 0 ? NF : 1];
};

template <int NF, class P>
__device__ __forceinline__ RtConstFb<NF> make_rtconst_fb(const P &p)
{
	RtConstFb<NF> r;
#pragma unroll
	for (int i = 0; i < (NF > 0 ? NF : 1); ++i) {
		/* a serving slot carries NF channels whatever its request asked
		 * for: the unused ones repeat channel 0 (their results are dropped) */
		const double nu = NF > 0 ? (double)p.nu[i < p.n_freq ? i : 0] : 1.0;
		r.c[i] = 1.0 / sqrt(nu);
		r.d[i] = 0.1 / (nu * nu * sqrt(nu));
	}
	return r;
}

/* The synthetic-disk transfer RHS (flow_disk + rt_rhs of gray_device.h)
 * from the Kerr-Schild quantities K: k_mu = eta u + f (l.u) l_mu and
 * g(v,v) = eta(v,v) + f (l.v)^2 for the Keplerian v. */
template <int NF, class P>
__device__ __forceinline__ void rt_rhs_fb(const P &p, const RtConstFb<NF> &rc, const V4<double> &q,
                                          const V4<double> &u, const KsAux<double> &K,
                                          const double *tau, double *dI, double *dtau)
{
	typedef MthFb M;
	double n = 0.0, sh = 1.0;
	if (p.fluid == GRAYHIP_FLUID_DISK) {
		const double R2 = q.x * q.x + q.y * q.y;
		if (K.r >= p.r_isco && R2 > 0.0) {
			const double isr = M::rsq(K.r);
			const double rs = (K.r * K.r) * isr;
			const double Om = M::rcp(rs + p.a);
			const double lv = 1.0 + Om * (q.x * K.ly - q.y * K.lx);
			const double vv = (Om * Om * R2 - 1.0) + K.f * lv * lv;
			if (vv < 0.0) {
				const double ut = M::rsq(-vv);
				const double flu = K.f * K.lu;
				const double kx = u.x + flu * K.lx, ky = u.y + flu * K.ly;
				const double kv = (flu - u.t) + Om * (q.x * ky - q.y * kx);
				const double s_ = -kv * ut;
				if (s_ > 0.0) {
					sh = s_;
					n = (K.ir * isr) * M::exp_(-(q.z * q.z) * M::rcp(0.18 * R2));
				}
			}
		}
	} else if (p.fluid == GRAYHIP_FLUID_UNIFORM) {
		n = 1.0;
	}
	if (n == 0.0) {
#pragma unroll
		for (int i = 0; i < NF; ++i) {
			dI[i] = 0.0;
			dtau[i] = 0.0;
		}
		return;
	}
	const double rsh = M::rsq(sh);
	const double A = n * rsh * M::rcp(sh * sh + 1e-28);
	const double Bc = n * (rsh * rsh) * rsh;
#pragma unroll
	for (int i = 0; i < NF; ++i) {
		dI[i] = -(A * rc.c[i]) * M::exp_(-tau[i]);
		dtau[i] = -(Bc * rc.d[i]);
	}
}

template <int NF, class P>
__device__ __forceinline__ void rhs_fb(const P &p, const RtConstFb<NF> &rc, const double *s,
                                       double *ds)
{
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	KsAux<double> K;
	V4<double> a = gr_rhs<double, MthFb>(p.a, q, u, NF > 0 ? &K : nullptr);
	ds[0] = u.t; ds[1] = u.x; ds[2] = u.y; ds[3] = u.z;
	ds[4] = a.t; ds[5] = a.x; ds[6] = a.y; ds[7] = a.z;
	if (NF > 0)
		rt_rhs_fb<NF, P>(p, rc, q, u, K, s + 8 + NF, ds + 8, ds + 8 + NF);
}

/* sim-org/RK4.cl:37-47 over rhs_fb. */
template <int NF, class P>
__device__ __forceinline__ void rk4_fb(const P &p, const RtConstFb<NF> &rc, double *s, double dt)
{
	constexpr int ND = 8 + 2 * NF;
	double k1[ND], k[ND], x[ND], k23[ND];
	const double hdt = 0.5 * dt;
	rhs_fb<NF, P>(p, rc, s, k1);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = s[i] + hdt * k1[i];
	rhs_fb<NF, P>(p, rc, x, k23);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = s[i] + hdt * k23[i];
	rhs_fb<NF, P>(p, rc, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		k1[i] = k1[i] + 2.0 * (k23[i] + k[i]);
		x[i] = s[i] + dt * k[i];
	}
	rhs_fb<NF, P>(p, rc, x, k);
	const double dt6 = dt * (1.0 / 6.0);
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = s[i] + dt6 * (k1[i] + k[i]);
}

/* The quantised sub-steps of one interval (as interval_f64) over rk4_fb. */
template <int NF, class P>
__device__ __forceinline__ int64_t interval_fb(const P &p, const RtConstFb<NF> &rc,
                                               const StepRule &sr, double *s, int h0,
                                               bool *retired)
{
	int64_t st = 0;
	int dh;
	*retired = false;
	for (int h = h0; h < sr.n; h += dh) {
		V4<double> q = {s[0], s[1], s[2], s[3]};
		V4<double> u = {s[4], s[5], s[6], s[7]};
		double x = getdt(p, q, u, sr.dts) / sr.dt * (double)sr.n;
		dh = (x == x) ? (int)x : 0;
		if (!dh) {
			*retired = true;
			break;
		}
		if (dh > sr.n - h)
			dh = sr.n - h;
		rk4_fb<NF, P>(p, rc, s, (double)dh * sr.dt / (double)sr.n);
		++st;
	}
	return st;
}

} // namespace gray
