/*
 * step_f32.h -- one ray in the FP32 form: FP32 state derivative and the
 * classical RK4 step on a compensated (hi + lo) state.  P is the parameter
 * block (the render's RayParams or per-lane serving parameters).
 */
#pragma once

#include "gray_device.h"

namespace gray {

template <int NF, class P>
__device__ __forceinline__ void rhs_f32(const P &p, const float *s, float *ds)
{
	V4<float> q = {s[0], s[1], s[2], s[3]};
	V4<float> u = {s[4], s[5], s[6], s[7]};
	V4<float> a = gr_rhs((float)p.a, q, u);
	ds[0] = u.t; ds[1] = u.x; ds[2] = u.y; ds[3] = u.z;
	ds[4] = a.t; ds[5] = a.x; ds[6] = a.y; ds[7] = a.z;
	if (NF > 0)
		rt_rhs<float, NF>(p, q, u, s + 8 + NF, ds + 8, ds + 8 + NF);
}

/* Classical RK4 (sim-org/RK4.cl:37-47) with stages taken from hi and the
 * increment added to (hi, lo) by compensated summation. */
template <int NF, class P>
__device__ __forceinline__ void rk4_f32c(const P &p, float *hi, float *lo, float dt)
{
	constexpr int ND = 8 + 2 * NF;
	float k[ND], acc[ND], x[ND];
	const float hdt = 0.5f * dt;

	rhs_f32<NF, P>(p, hi, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] = k[i];
		x[i] = hi[i] + hdt * k[i];
	}
	rhs_f32<NF, P>(p, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + hdt * k[i];
	}
	rhs_f32<NF, P>(p, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + dt * k[i];
	}
	rhs_f32<NF, P>(p, x, k);
	const float dt6 = dt * (1.0f / 6.0f);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		float y = (acc[i] + k[i]) * dt6 + lo[i];
		float t = hi[i] + y;
		lo[i] = y - (t - hi[i]);
		hi[i] = t;
	}
}

template <int NF>
__device__ __forceinline__ void store_state(double *__restrict__ data, size_t n_rays, size_t g,
                                            const float *hi, const float *lo)
{
	constexpr int ND = 8 + 2 * NF;
#pragma unroll
	for (int i = 0; i < ND; ++i)
		data[(size_t)i * n_rays + g] = (double)hi[i] + (double)lo[i];
}

} // namespace gray
