/*
 * step_fb.h -- the FP64 step of the FP32 form's re-integration (rays whose
 * FP32 step collapses, handed over by a pool or a serving epoch).
 *
 * These rays are not part of the bit-exact form, and their dependency chain
 * of sequential steps is the critical path of a render's tail, so the step
 * is built for few instructions per step at FP64 accuracy:
 *   - divisions and square roots from the hardware FP64 reciprocal and
 *     reciprocal square root refined by two Newton steps (MthFb of
 *     rt_disk.h: about 1 ulp instead of correctly rounded);
 *   - the transfer RHS of rt_disk.h, the channels' depths as one shared
 *     integral T (the state is nd64 reals, as the plain form's);
 *   - the quantised step count as getdt (n / dt), one multiply;
 *   - the RK4 combination multiplied by 1/6.
 * The arithmetic is the same mathematics as sim-org/KS.cl:263-378,
 * rt.cl:185-198 and RK4.cl:37-47 in a different evaluation order.
 */
#pragma once

#include "gray_device.h"
#include "rt_disk.h"
#include "step_f64.h"

namespace gray {

template <int NF, class P>
__device__ __forceinline__ void rhs_fb(const P &p, const RtConst64<NF> &rc, const double *s,
                                       double *ds)
{
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	KsAux<double> K;
	V4<double> a = gr_rhs<double, MthFb>(p.a, q, u, NF > 0 ? &K : nullptr);
	ds[0] = u.t; ds[1] = u.x; ds[2] = u.y; ds[3] = u.z;
	ds[4] = a.t; ds[5] = a.x; ds[6] = a.y; ds[7] = a.z;
	if constexpr (NF > 0)
		rt_rhs_disk64_shared<NF, P>(p, rc, q, u, K, s[8 + NF], ds + 8, ds[8 + NF]);
}

/* sim-org/RK4.cl:37-47 over rhs_fb. */
template <int NF, class P>
__device__ __forceinline__ void rk4_fb(const P &p, const RtConst64<NF> &rc, double *s, double dt)
{
	constexpr int ND = nd64<NF>();
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
__device__ __forceinline__ int64_t interval_fb(const P &p, const RtConst64<NF> &rc,
                                               const StepRule &sr, double *s, int h0,
                                               bool *retired)
{
	int64_t st = 0;
	int dh;
	*retired = false;
	const double n_over_dt = (double)sr.n / sr.dt;
	for (int h = h0; h < sr.n; h += dh) {
		V4<double> q = {s[0], s[1], s[2], s[3]};
		V4<double> u = {s[4], s[5], s[6], s[7]};
		double x = getdt(p, q, u, sr.dts) * n_over_dt;
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
