/*
 * grid_f64.h -- FP64 interpolation of a resident spacetime: per-field
 * trilinear sums in corner order, time blend and g(u,u), in the evaluation
 * order of the FP64 grid oracle (the dyst RHS over them is in
 * kernels_grid64.hip).
 */
#pragma once

#include "grid_device.h"
#include "rt_thermal.h"

namespace gray {

/* Fields base..base+NC-1 of one level (interleaved, stride NCOMP) summed over
 * the 8 corners in corner order: s = w0 v0; s = s + wk vk. */
template <int NC, int NCOMP, class S>
__device__ __forceinline__ void sample_lev_f64(const S *__restrict__ a, int base,
                                               const Stencil<double> &st, double out[NC])
{
	{
		const S *p = a + (size_t)st.corner(0) * NCOMP + base;
		const double w = st.weight(0);
#pragma unroll
		for (int c = 0; c < NC; ++c)
			out[c] = w * (double)p[c];
	}
#pragma unroll 1
	for (int k = 1; k < 8; ++k) {
		const S *p = a + (size_t)st.corner(k) * NCOMP + base;
		const double w = st.weight(k);
#pragma unroll
		for (int c = 0; c < NC; ++c)
			out[c] = out[c] + w * (double)p[c];
	}
}

/* Both levels blended in time (sim-org/interp.cl:183-205). */
template <int NC, int NCOMP, class S>
__device__ __forceinline__ void sample_f64(const S *const lev[2], int base,
                                           const Stencil<double> &st, double out[NC])
{
	sample_lev_f64<NC, NCOMP, S>(lev[0], base, st, out);
	/* wt = 0 (ray time clamped to t1) reads only the t1 level: for finite
	 * data y1 + 0 (y2 - y1) = y1, so skipping the second level is exact */
	if (st.wt > 0.0) {
		double y2[NC];
		sample_lev_f64<NC, NCOMP, S>(lev[1], base, st, y2);
#pragma unroll
		for (int c = 0; c < NC; ++c)
			out[c] = out[c] + st.wt * (y2[c] - out[c]);
	}
}

/* k_mu = g_{mu nu} u^nu with the interpolated metric (sim-org/dyst.cl:55-81,
 * with g_zz read from g_zz); s = (q, u). */
template <class S>
__device__ __forceinline__ void grid_down_f64(const GridView<S> &G, const double *s, double k[4])
{
	Stencil<double> st;
	make_stencil<double, S>(G, s[0], s[1], s[2], s[3], st);
	double m[10];
	sample_f64<10, 10, S>(G.met, 0, st, m);
	const double u[4] = {s[4], s[5], s[6], s[7]};
#pragma unroll
	for (int a = 0; a < 4; ++a)
		k[a] = ((m[sym(a, 0)] * u[0] + m[sym(a, 1)] * u[1]) + m[sym(a, 2)] * u[2]) +
		       m[sym(a, 3)] * u[3];
}

template <class S>
__device__ __forceinline__ double grid_getuu_f64(const GridView<S> &G, const double *s)
{
	double k[4];
	grid_down_f64<S>(G, s, k);
	return ((s[4] * k[0] + s[5] * k[1]) + s[6] * k[2]) + s[7] * k[3];
}

} // namespace gray
