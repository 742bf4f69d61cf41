/*
 * grid_f32.h -- FP32 form of the gridded RHS: per corner, contract the 40
 * Christoffels with the ten products u^a u^b, then interpolate the four
 * results; RK4 on a compensated (hi + lo) state.
 */
#pragma once

#include "grid_device.h"

namespace gray {

/* 40 values of one grid point as floats, read with 16-byte loads. */
__device__ __forceinline__ void load40(const float *__restrict__ a, float v[40])
{
	const float4 *a4 = reinterpret_cast<const float4 *>(a);
#pragma unroll
	for (int i = 0; i < 10; ++i) {
		float4 t = a4[i];
		v[4 * i] = t.x;
		v[4 * i + 1] = t.y;
		v[4 * i + 2] = t.z;
		v[4 * i + 3] = t.w;
	}
}

__device__ __forceinline__ void load40(const double *__restrict__ a, float v[40])
{
	const double2 *a2 = reinterpret_cast<const double2 *>(a);
#pragma unroll
	for (int i = 0; i < 20; ++i) {
		double2 t = a2[i];
		v[2 * i] = (float)t.x;
		v[2 * i + 1] = (float)t.y;
	}
}

/* sum_k w_k (Gamma^mu_ab(corner k) P_ab) for one level, mu = 0..3 */
template <class S>
__device__ __forceinline__ void lev_contract(const S *__restrict__ lev, const Stencil<float> &st,
                                             const float P[10], float out[4])
{
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
		out[mu] = 0.f;
#pragma unroll 1
	for (int k = 0; k < 8; ++k) {
		float v[40];
		load40(lev + (size_t)st.corner(k) * 40, v);
		const float w = st.weight(k);
#pragma unroll
		for (int mu = 0; mu < 4; ++mu) {
			float t = 0.f;
#pragma unroll
			for (int c = 0; c < 10; ++c)
				t += v[10 * mu + c] * P[c];
			out[mu] += w * t;
		}
	}
}

template <class S>
__device__ __forceinline__ void grid_rhs_f32(const GridView<S> &G, const float *s, float *du)
{
	Stencil<float> st;
	make_stencil<float, S>(G, s[0], s[1], s[2], s[3], st);
	const float u0 = s[4], u1 = s[5], u2 = s[6], u3 = s[7];
	const float P[10] = {u0 * u0, 2.f * u0 * u1, 2.f * u0 * u2, 2.f * u0 * u3, u1 * u1,
	                     2.f * u1 * u2, 2.f * u1 * u3, u2 * u2, 2.f * u2 * u3, u3 * u3};
	float uGu[4];
	lev_contract<S>(G.gam[0], st, P, uGu);
	if (st.wt >= 0.f) {
		float y2[4];
		lev_contract<S>(G.gam[1], st, P, y2);
#pragma unroll
		for (int mu = 0; mu < 4; ++mu)
			uGu[mu] = uGu[mu] + st.wt * (y2[mu] - uGu[mu]);
	}
	const float u[4] = {u0, u1, u2, u3};
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
		du[mu] = -uGu[mu] + uGu[0] * u[mu];
}

/* Classical RK4 with stages from hi, increment added to (hi, lo). */
template <class S>
__device__ __forceinline__ void grid_rk4_f32c(const GridView<S> &G, float *hi, float *lo, float dt)
{
	float k[8], acc[8], x[8];
	const float hdt = 0.5f * dt;
#define GRHS(in, out) do { out[0] = in[4]; out[1] = in[5]; out[2] = in[6]; out[3] = in[7]; \
		grid_rhs_f32<S>(G, in, out + 4); } while (0)
	GRHS(hi, k);
#pragma unroll
	for (int i = 0; i < 8; ++i) {
		acc[i] = k[i];
		x[i] = hi[i] + hdt * k[i];
	}
	GRHS(x, k);
#pragma unroll
	for (int i = 0; i < 8; ++i) {
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + hdt * k[i];
	}
	GRHS(x, k);
#pragma unroll
	for (int i = 0; i < 8; ++i) {
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + dt * k[i];
	}
	GRHS(x, k);
#undef GRHS
	const float dt6 = dt * (1.0f / 6.0f);
#pragma unroll
	for (int i = 0; i < 8; ++i) {
		float y = (acc[i] + k[i]) * dt6 + lo[i];
		float t = hi[i] + y;
		lo[i] = y - (t - hi[i]);
		hi[i] = t;
	}
}

} // namespace gray
