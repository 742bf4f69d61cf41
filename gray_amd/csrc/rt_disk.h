/*
 * rt_disk.h -- the FP64 transfer RHS of the synthetic disk, evaluated from
 * the Kerr-Schild quantities the geodesic RHS computes anyway, and the FP64
 * reciprocal and reciprocal square root it runs on (~4e-15 relative).
 *
 * Used by the plain FP64 form (whose geodesic stays bit-exact: the transfer
 * channels never feed back into the geodesic, and they are held to 1e-10
 * relative, exp not being correctly rounded in any case) and by the FP64
 * re-integration of the FP32 form.  Same mathematics as sim-org/rt.cl:185-198
 * with the fluid of sim-org/flow.cl:36-58 replaced by the Keplerian disk,
 * regrouped:
 *   k_mu = eta u + f (l.u) l_mu,   g(v,v) = eta(v,v) + f (l.v)^2
 * for the Keplerian v = (1, -y Omega, x Omega, 0), Omega = 1/(r^1.5 + a);
 * s = -k.v / sqrt(-g(v,v)), n = r^-1.5 exp(-z^2/(2 (0.3 R)^2)) outside
 * r_isco, and per channel
 *   dI_i = -n nu_i^-1/2 s^-1/2 exp(-tau_i) / (s^2 + 1e-28)
 *   dtau_i = -0.1 n nu_i^-5/2 s^-3/2
 * with nu_i^-1/2 and 0.1 nu_i^-5/2 formed once per ray.
 */
#pragma once

#include "gray_device.h"

namespace gray {

/* FP64 reciprocal, reciprocal square root, division and square root for
 * the non-bitwise FP64 code (transfer channels, the FP32 form's
 * re-integration): the hardware estimates, within 5.2e-8 relative on
 * gfx950 (measured over 2^20 doubles in [2^-30, 2^30]), refined by one
 * Newton step, which squares that: within ~4e-15 relative. */
struct MthFb {
	static __device__ __forceinline__ double rcp(double b)
	{
		const double r = __builtin_amdgcn_rcp(b);
		return __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
	}
	static __device__ __forceinline__ double rsq(double x)
	{
		const double y = __builtin_amdgcn_rsq(x);
		/* y (3 - x y^2) / 2 */
		return y * __builtin_fma(-0.5 * x * y, y, 1.5);
	}
	static __device__ __forceinline__ double div(double a, double b) { return a * rcp(b); }
	static __device__ __forceinline__ double sqrt_(double x)
	{
		return x > 0.0 ? x * rsq(x) : __builtin_sqrt(x);
	}
	static __device__ __forceinline__ double exp_(double x) { return exp(x); }
};

template <int NF> struct RtConst64 {
	double c[NF > 0 ? NF : 1], d[NF > 0 ? NF : 1];
};

template <int NF, class P>
__device__ __forceinline__ RtConst64<NF> make_rtconst64(const P &p)
{
	RtConst64<NF> r;
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

/* k_mu = eta_mu nu u^nu + f (l.u) l_mu from the Kerr-Schild quantities of
 * gr_rhs at (q, u) (the u^mu forms; the Hamiltonian forms carry k_mu). */
__device__ __forceinline__ V4<double> cov_from_aux(const V4<double> &u, const KsAux<double> &K)
{
	const double flu = K.f * K.lu;
	return V4<double>{flu - u.t, u.x + flu * K.lx, u.y + flu * K.ly, u.z + flu * K.lz};
}

/* Density n and redshift factor sh of the disk at q for the covariant
 * photon momentum k; n = 0 where nothing is emitted.  K must come from the
 * geodesic RHS at q; its r is the r of getrr, so the r_isco cut is the one
 * of the step rule's geometry. */
template <class P>
__device__ __forceinline__ void disk_emission64(const P &p, const V4<double> &q,
                                                const V4<double> &k, const KsAux<double> &K,
                                                double &n, double &sh)
{
	typedef MthFb M;
	n = 0.0;
	sh = 1.0;
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
				const double kv = k.t + Om * (q.x * k.y - q.y * k.x);
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
}

/* exp(-d_i T) of the channels at the shared depth T0 of an RK4 step's first
 * stage; the later stages' factors follow from them (chan_exp).  E is
 * filled at once (chan_exp_init) or at the first stage whose emission is
 * non-zero (chan_exp_defer; same values), which spares the exponentials to
 * steps that emit nothing, such as a ray's approach to the horizon inside
 * r_isco. */
template <int NF> struct ChanExp {
	double E[NF > 0 ? NF : 1];
	double T0;
	bool ready;   /* E holds exp(-d_i T0) */
};

template <int NF>
__device__ __forceinline__ void chan_exp_fill(const RtConst64<NF> &rc, ChanExp<NF> &ce)
{
#pragma unroll
	for (int i = 0; i < NF; ++i)
		ce.E[i] = MthFb::exp_(-(rc.d[i] * ce.T0));
	ce.ready = true;
}

template <int NF>
__device__ __forceinline__ ChanExp<NF> chan_exp_init(const RtConst64<NF> &rc, double T0)
{
	ChanExp<NF> ce;
	ce.T0 = T0;
	chan_exp_fill<NF>(rc, ce);
	return ce;
}

template <int NF>
__device__ __forceinline__ ChanExp<NF> chan_exp_defer(double T0)
{
	ChanExp<NF> ce;
	ce.T0 = T0;
#pragma unroll
	for (int i = 0; i < (NF > 0 ? NF : 1); ++i)
		ce.E[i] = 0.0;
	ce.ready = false;
	return ce;
}

/* exp(-d_i T) for every channel at a stage's depth T: exp(-d_i T0) times
 * exp(x_i), x_i = -d_i (T - T0), by its Taylor series to degree 5 where
 * every |x_i| <= 2^-10 (truncation below 2^-69 relative; T moves little
 * within a step), the exponential itself otherwise. */
template <int NF>
__device__ __forceinline__ void chan_exp(const RtConst64<NF> &rc, ChanExp<NF> &ce, double T,
                                         double *e)
{
	if (!ce.ready)
		chan_exp_fill<NF>(rc, ce);
	const double dT = T - ce.T0;
	bool small = true;
	double x[NF > 0 ? NF : 1];
#pragma unroll
	for (int i = 0; i < NF; ++i) {
		x[i] = -(rc.d[i] * dT);
		small = small && fabs(x[i]) <= 0x1p-10;
	}
	if (small) {
#pragma unroll
		for (int i = 0; i < NF; ++i) {
			double t = __builtin_fma(x[i], 1.0 / 120.0, 1.0 / 24.0);
			t = __builtin_fma(x[i], t, 1.0 / 6.0);
			t = __builtin_fma(x[i], t, 0.5);
			t = __builtin_fma(x[i], t, 1.0);
			e[i] = ce.E[i] * __builtin_fma(x[i], t, 1.0);
		}
	} else {
#pragma unroll
		for (int i = 0; i < NF; ++i)
			e[i] = MthFb::exp_(-(rc.d[i] * T));
	}
}

/* The transfer RHS (sim-org/rt.cl:185-198) with the channels' depths as one
 * shared integral: the disk's absorption is a fixed power of frequency, so
 * tau_i = d_i T with d_i = 0.1 nu_i^-5/2 (rc.d) and dT = -n sh^-3/2.  k is
 * the covariant photon momentum, ce the channels' exponentials at the
 * step's first stage.  Fills dI[0..NF), dT. */
template <int NF, class P>
__device__ __forceinline__ void rt_rhs_disk64_shared(const P &p, const RtConst64<NF> &rc,
                                                     ChanExp<NF> &ce,
                                                     const V4<double> &q, const V4<double> &k,
                                                     const KsAux<double> &K, double T,
                                                     double *dI, double &dT)
{
	typedef MthFb M;
	double n, sh;
	disk_emission64(p, q, k, K, n, sh);
	if (n == 0.0) {
#pragma unroll
		for (int i = 0; i < NF; ++i)
			dI[i] = 0.0;
		dT = 0.0;
		return;
	}
	const double rsh = M::rsq(sh);
	const double A = n * rsh * M::rcp(sh * sh + 1e-28);
	dT = -(n * (rsh * rsh) * rsh);
	double e[NF > 0 ? NF : 1];
	chan_exp<NF>(rc, ce, T, e);
#pragma unroll
	for (int i = 0; i < NF; ++i)
		dI[i] = -(A * rc.c[i]) * e[i];
}

} // namespace gray
