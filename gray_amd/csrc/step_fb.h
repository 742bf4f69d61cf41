/*
 * step_fb.h -- the FP64 step of the FP32 form's re-integration (rays whose
 * FP32 step collapses, handed over by a pool or a serving epoch).
 *
 * These rays are not part of the bit-exact form, and their dependency chain
 * of sequential steps is the critical path of a render's tail, so the step
 * is built for few instructions per step at FP64 accuracy:
 *   - the Hamiltonian form of the geodesic, as the FP32 form's (ks_ham:
 *     state t, x, y, z, p_mu, p_t conserved), with the reciprocals and
 *     reciprocal square roots from the hardware FP64 instructions refined
 *     by a Newton step (MthFb of rt_disk.h: ~4e-15 relative instead of
 *     correctly rounded);
 *   - the transfer RHS of rt_disk.h on k_mu = p_mu, the channels' depths as
 *     one shared integral T (the state is nd64 reals, as the plain form's),
 *     their exponentials from the step's first stage (chan_exp), formed
 *     only in steps that emit;
 *   - the quantised step count as getdt (n / dt), one multiply, with u^mu
 *     and r from the step's first RK4 stage;
 *   - the RK4 combination multiplied by 1/6.
 * The arithmetic is the same mathematics as sim-org/KS.cl:263-378,
 * rt.cl:185-198 and RK4.cl:37-47 in a different form and evaluation order.
 * fb_load / fb_store convert from and to the stored layout (u^mu) in FP64.
 */
#pragma once

#include "gray_device.h"
#include "rt_disk.h"
#include "step_f64.h"

namespace gray {

template <int NF, class P>
__device__ __forceinline__ void rhs_fb(const P &p, const RtConst64<NF> &rc,
                                       ChanExp<NF> &ce, const double *s, double *ds,
                                       double *r_out = nullptr)
{
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> k = {s[4], s[5], s[6], s[7]};
	KsAux<double> K;
	const Ham<double> h = ks_ham<double, MthFb>(p.a, q, k, NF > 0 ? &K : nullptr);
	ds[0] = h.u.t; ds[1] = h.u.x; ds[2] = h.u.y; ds[3] = h.u.z;
	ds[4] = 0.0;   ds[5] = h.dpx; ds[6] = h.dpy; ds[7] = h.dpz;
	if (r_out)
		*r_out = h.r;
	if constexpr (NF > 0)
		rt_rhs_disk64_shared<NF, P>(p, rc, ce, q, k, K, s[8 + NF], ds + 8, ds[8 + NF]);
}

/* sim-org/RK4.cl:37-47 over rhs_fb; k1 is rhs_fb at s, which the caller has
 * evaluated. */
template <int NF, class P>
__device__ __forceinline__ void rk4_fb(const P &p, const RtConst64<NF> &rc,
                                       ChanExp<NF> &ce, double *s, const double *k1in,
                                       double dt)
{
	constexpr int ND = nd64<NF>();
	double k1[ND], k[ND], x[ND], k23[ND];
	const double hdt = 0.5 * dt;
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		k1[i] = k1in[i];
		x[i] = s[i] + hdt * k1[i];
	}
	rhs_fb<NF, P>(p, rc, ce, x, k23);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = s[i] + hdt * k23[i];
	rhs_fb<NF, P>(p, rc, ce, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		k1[i] = k1[i] + 2.0 * (k23[i] + k[i]);
		x[i] = s[i] + dt * k[i];
	}
	rhs_fb<NF, P>(p, rc, ce, x, k);
	const double dt6 = dt * (1.0 / 6.0);
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = s[i] + dt6 * (k1[i] + k[i]);
}

/* One quantised sub-step (sim-org/driver.cl:81-99) over rk4_fb: returns the
 * units dh it advanced, 0 where the step rule stops the ray. */
template <int NF, class P>
__device__ __forceinline__ int step_fb(const P &p, const RtConst64<NF> &rc, const StepRule &sr,
                                       double *s, int h)
{
	constexpr int ND = nd64<NF>();
	double k1[ND];
	ChanExp<NF> ce = chan_exp_defer<NF>(NF > 0 ? s[ND - 1] : 0.0);
	double r1;
	rhs_fb<NF, P>(p, rc, ce, s, k1, &r1);
	const V4<double> q = {s[0], s[1], s[2], s[3]};
	const V4<double> u = {k1[0], k1[1], k1[2], k1[3]};
	const double x = getdt_r(p, r1, q, u, sr.dts) * ((double)sr.n / sr.dt);
	int dh = (x == x) ? (int)x : 0;
	if (!dh)
		return 0;
	if (dh > sr.n - h)
		dh = sr.n - h;
	rk4_fb<NF, P>(p, rc, ce, s, k1, (double)dh * sr.dt / (double)sr.n);
	return dh;
}

/* The quantised sub-steps of one interval (as interval_f64) over step_fb. */
template <int NF, class P>
__device__ __forceinline__ int64_t interval_fb(const P &p, const RtConst64<NF> &rc,
                                               const StepRule &sr, double *s, int h0,
                                               bool *retired)
{
	int64_t st = 0;
	int dh;
	*retired = false;
	for (int h = h0; h < sr.n; h += dh) {
		dh = step_fb<NF, P>(p, rc, sr, s, h);
		if (!dh) {
			*retired = true;
			break;
		}
		++st;
	}
	return st;
}

/* Stored geodesic (t, x, y, z, u^mu) -> the step's (t, x, y, z, p_mu), in
 * place, and back. */
__device__ __forceinline__ void fb_lower(double a, double *s)
{
	const V4<double> k = down(a, V4<double>{s[0], s[1], s[2], s[3]},
	                          V4<double>{s[4], s[5], s[6], s[7]});
	s[4] = k.t; s[5] = k.x; s[6] = k.y; s[7] = k.z;
}

__device__ __forceinline__ void fb_raise(double a, double *s)
{
	const V4<double> u = up(a, V4<double>{s[0], s[1], s[2], s[3]},
	                        V4<double>{s[4], s[5], s[6], s[7]});
	s[4] = u.t; s[5] = u.x; s[6] = u.y; s[7] = u.z;
}

} // namespace gray
