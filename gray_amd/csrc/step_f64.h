/*
 * step_f64.h -- one ray in FP64: state derivative, classical RK4 step, the
 * quantised sub-steps of one dump interval, and the camera initial state.
 * Included by the bit-exact FP64 kernels (compiled without contraction)
 * and by the FP64 re-integration of the FP32 form (compiled with it).
 */
#pragma once

#include "gray_device.h"
#include "rt_disk.h"

namespace gray {

/* Reals the plain FP64 form integrates: geodesic, intensities and the
 * channels' shared depth integral T (tau_i = d_i T, rt_disk.h). */
template <int NF> constexpr int nd64() { return 8 + NF + (NF > 0 ? 1 : 0); }

/* The geodesic in the written order with correctly rounded division and
 * square root (bit-exact); the transfer channels from rt_disk.h. */
template <int NF, class P>
__device__ __forceinline__ void rhs_f64(const P &p, const RtConst64<NF> &rc,
                                        ChanExp<NF> &ce, const double *s, double *ds)
{
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	KsAux<double> K;
	V4<double> a = gr_rhs(p.a, q, u, NF > 0 ? &K : nullptr);
	ds[0] = u.t; ds[1] = u.x; ds[2] = u.y; ds[3] = u.z;
	ds[4] = a.t; ds[5] = a.x; ds[6] = a.y; ds[7] = a.z;
	if constexpr (NF > 0)
		rt_rhs_disk64_shared<NF, P>(p, rc, ce, q, cov_from_aux(u, K), K, s[8 + NF], ds + 8,
		                            ds[8 + NF]);
}

/* sim-org/RK4.cl:37-47: s + dt (k1 + 2 (k2 + k3) + k4) / 6, stages at
 * s + (dt/2) k1, s + (dt/2) k2, s + dt k3, each stage a fused multiply-add
 * (as the geodesic RHS, gray_device.h); k1 + 2 (k2 + k3) likewise. */
template <int NF, class P>
__device__ __forceinline__ void rk4_f64(const P &p, const RtConst64<NF> &rc, double *s, double dt)
{
	constexpr int ND = nd64<NF>();
	double k1[ND], k[ND], x[ND], k23[ND];
	const double hdt = 0.5 * dt;
	ChanExp<NF> ce = chan_exp_init<NF>(rc, NF > 0 ? s[ND - 1] : 0.0);

	rhs_f64<NF, P>(p, rc, ce, s, k1);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = fmaT(hdt, k1[i], s[i]);
	rhs_f64<NF, P>(p, rc, ce, x, k23);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = fmaT(hdt, k23[i], s[i]);
	rhs_f64<NF, P>(p, rc, ce, x, k);
	/* k1 + 2 (k2 + k3) is formed before the last stage (same value), so
	 * k1 and k2 + k3 are not both live across it */
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		k1[i] = fmaT(2.0, k23[i] + k[i], k1[i]);
		x[i] = fmaT(dt, k[i], s[i]);
	}
	rhs_f64<NF, P>(p, rc, ce, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = s[i] + div6(dt * (k1[i] + k[i]));
}

/* Sub-steps of one interval from unit h0 (sim-org/driver.cl:92-99).
 * Returns steps taken; *retired set when the step rule returned zero. */
template <int NF, class P>
__device__ __forceinline__ int64_t interval_f64(const P &p, const RtConst64<NF> &rc,
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
		rk4_f64<NF, P>(p, rc, s, (double)dh * sr.dt / (double)sr.n);
		++st;
	}
	return st;
}

template <int NF, class P>
__device__ __forceinline__ void icond_ray(const P &p, size_t g, double *s)
{
	size_t gi = g % (size_t)p.w_rays, gj = g / (size_t)p.w_rays;
	double alpha = (((double)gi + 0.5) / (double)p.w_rays - 0.5) * p.w_img;
	double beta  = (((double)gj + 0.5) / (double)p.h_rays - 0.5) * p.h_img;
	V4<double> q, u;
	gr_icond(p.a, p.r_obs, p.si, p.ci, p.sj, p.cj, alpha, beta, q, u);
	s[0] = q.t; s[1] = q.x; s[2] = q.y; s[3] = q.z;
	s[4] = u.t; s[5] = u.x; s[6] = u.y; s[7] = u.z;
#pragma unroll
	for (int i = 0; i < 2 * NF; ++i)
		s[8 + i] = 0.0;
}

} // namespace gray
