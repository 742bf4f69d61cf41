/*
 * grid_f32.h -- FP32 form of the gridded RHS: per corner, contract the 40
 * Christoffels with the ten products u^a u^b, then interpolate the four
 * results; RK4 on a compensated (hi + lo) state.  With transfer, the
 * metric row g_{t nu} and the density are contracted and interpolated the
 * same way.
 */
#pragma once

#include "grid_device.h"
#include "rt_thermal.h"

namespace gray {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

/* 40 values of one grid point as floats, read with 16-byte loads. */
__device__ __forceinline__ void load40(gptr<float> a, float v[40])
{
	gptr<f32x4> a4 = (gptr<f32x4>)a;
#pragma unroll
	for (int i = 0; i < 10; ++i) {
		const f32x4 t = a4[i];
		v[4 * i] = t.x;
		v[4 * i + 1] = t.y;
		v[4 * i + 2] = t.z;
		v[4 * i + 3] = t.w;
	}
}

__device__ __forceinline__ void load40(gptr<double> a, float v[40])
{
	gptr<f64x2> a2 = (gptr<f64x2>)a;
#pragma unroll
	for (int i = 0; i < 20; ++i) {
		const f64x2 t = a2[i];
		v[2 * i] = (float)t.x;
		v[2 * i + 1] = (float)t.y;
	}
}

/* Half a point's Christoffels (mu = 0, 1 or mu = 2, 3): a corner is read
 * in two halves so that only 20 values are live at a time. */
__device__ __forceinline__ void load20(gptr<float> a, float v[20])
{
	gptr<f32x4> a4 = (gptr<f32x4>)a;
#pragma unroll
	for (int i = 0; i < 5; ++i) {
		const f32x4 t = a4[i];
		v[4 * i] = t.x;
		v[4 * i + 1] = t.y;
		v[4 * i + 2] = t.z;
		v[4 * i + 3] = t.w;
	}
}

__device__ __forceinline__ void load20(gptr<double> a, float v[20])
{
	gptr<f64x2> a2 = (gptr<f64x2>)a;
#pragma unroll
	for (int i = 0; i < 10; ++i) {
		const f64x2 t = a2[i];
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
		const float w = st.weight(k);
#pragma unroll
		for (int hf = 0; hf < 2; ++hf) {
			float v[20];
			load20(gmem(lev) + (size_t)st.corner(k) * 40 + 20 * hf, v);
#pragma unroll
			for (int m2 = 0; m2 < 2; ++m2) {
				float t = 0.f;
#pragma unroll
				for (int c = 0; c < 10; ++c)
					t += v[10 * m2 + c] * P[c];
				out[2 * hf + m2] += w * t;
			}
		}
	}
}

/* lev_contract with the next half-corner's five 16-byte loads issued before
 * the current half is summed (FP32 storage; the same sums in the same
 * order); for the kernels that have the registers. */
template <class S>
__device__ __forceinline__ void lev_contract_pf(const S *__restrict__ lev, const Stencil<float> &st,
                                                const float P[10], float out[4])
{
	if constexpr (sizeof(S) != 4) {
		lev_contract<S>(lev, st, P, out);
	} else {
		gptr<f32x4> a4 = (gptr<f32x4>)gmem(lev);
		/* sixteen half-corners of 20 values; half j + 1's five loads are
		 * issued before half j is summed */
		f32x4 cur[5], nxt[5];
#pragma unroll
		for (int i = 0; i < 5; ++i)
			cur[i] = a4[(size_t)st.corner(0) * 10 + i];
#pragma unroll
		for (int mu = 0; mu < 4; ++mu)
			out[mu] = 0.f;
#pragma unroll
		for (int j = 0; j < 16; ++j) {
			const int k = j >> 1, hf = j & 1;
			if (j < 15) {
				const int k1 = (j + 1) >> 1, h1 = (j + 1) & 1;
#pragma unroll
				for (int i = 0; i < 5; ++i)
					nxt[i] = a4[(size_t)st.corner(k1) * 10 + 5 * h1 + i];
			}
			const float w = st.weight(k);
			float v[20];
#pragma unroll
			for (int i = 0; i < 5; ++i) {
				const f32x4 t = cur[i];
				v[4 * i] = t.x;
				v[4 * i + 1] = t.y;
				v[4 * i + 2] = t.z;
				v[4 * i + 3] = t.w;
			}
#pragma unroll
			for (int m2 = 0; m2 < 2; ++m2) {
				float t = 0.f;
#pragma unroll
				for (int c = 0; c < 10; ++c)
					t += v[10 * m2 + c] * P[c];
				out[2 * hf + m2] += w * t;
			}
#pragma unroll
			for (int i = 0; i < 5; ++i)
				cur[i] = nxt[i];
		}
	}
}

/* sum_k w_k (g_{t nu}(corner k) u^nu) and sum_k w_k rho(corner k): k_t and
 * n_e of one level (interpolation commutes with the contraction). */
__device__ __forceinline__ void load4(gptr<float> a, float v[4])
{
	gptr<f32x2> a2 = (gptr<f32x2>)a;   /* p*40 B: 8-byte aligned */
	const f32x2 x = a2[0], y = a2[1];
	v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
}

__device__ __forceinline__ void load4(gptr<double> a, float v[4])
{
	gptr<f64x2> a2 = (gptr<f64x2>)a;
	const f64x2 x = a2[0], y = a2[1];
	v[0] = (float)x.x; v[1] = (float)x.y; v[2] = (float)y.x; v[3] = (float)y.y;
}

template <class S>
__device__ __forceinline__ void lev_flow(const S *__restrict__ met, const S *__restrict__ rho,
                                         const Stencil<float> &st, const float u[4], float &kt,
                                         float &ne)
{
	kt = 0.f;
	ne = 0.f;
#pragma unroll 1
	for (int k = 0; k < 8; ++k) {
		const uint32_t p = st.corner(k);
		float m[4];
		load4(gmem(met) + (size_t)p * 10, m);
		const float w = st.weight(k);
		kt += w * (((m[0] * u[0] + m[1] * u[1]) + m[2] * u[2]) + m[3] * u[3]);
		ne += w * (float)gmem(rho)[p];
	}
}

/* lev_flow plus sum_k w_k (g_{z nu}(corner k) u^nu): k_z, from the z row
 * {3, 6, 8, 9} of the metric, for the field's bkcos. */
template <class S>
__device__ __forceinline__ void lev_flow_z(const S *__restrict__ met, const S *__restrict__ rho,
                                           const Stencil<float> &st, const float u[4], float &kt,
                                           float &kz, float &ne)
{
	kt = 0.f;
	kz = 0.f;
	ne = 0.f;
#pragma unroll 1
	for (int k = 0; k < 8; ++k) {
		const uint32_t p = st.corner(k);
		float m[4], mz[4];
		load4(gmem(met) + (size_t)p * 10, m);
		load4(gmem(met) + (size_t)p * 10 + 6, mz);
		const float w = st.weight(k);
		kt += w * (((m[0] * u[0] + m[1] * u[1]) + m[2] * u[2]) + m[3] * u[3]);
		kz += w * (((m[3] * u[0] + mz[0] * u[1]) + mz[2] * u[2]) + mz[3] * u[3]);
		ne += w * (float)gmem(rho)[p];
	}
}

/* k_t, n_e and bkcos (flow.cl:36-58, field b^mu = (0, 0, 0, B)) of both
 * levels blended in time; bkcos = 0 without a field. */
template <class S>
__device__ __forceinline__ void flow_f32(const GridView<S> &G, const RtThermalF &rt,
                                         const Stencil<float> &st, const float u[4], float &kt,
                                         float &ne, float &bkcos)
{
	bkcos = 0.f;
	if (rt.syn_on) {
		float kz;
		lev_flow_z<S>(G.met[0], G.rho[0], st, u, kt, kz, ne);
		if (st.wt > 0.f) {
			float kt2, kz2, ne2;
			lev_flow_z<S>(G.met[1], G.rho[1], st, u, kt2, kz2, ne2);
			kt = kt + st.wt * (kt2 - kt);
			kz = kz + st.wt * (kz2 - kz);
			ne = ne + st.wt * (ne2 - ne);
		}
		bkcos = (kz * rt.b) / ((-kt) * rt.b + 1e-28f);
		return;
	}
	lev_flow<S>(G.met[0], G.rho[0], st, u, kt, ne);
	if (st.wt > 0.f) {
		float kt2, ne2;
		lev_flow<S>(G.met[1], G.rho[1], st, u, kt2, ne2);
		kt = kt + st.wt * (kt2 - kt);
		ne = ne + st.wt * (ne2 - ne);
	}
}

/* Geodesic RHS and, for NF > 0, the transfer of the reference flow on the
 * scaled state (I' = I / K, tau' = tau / K). */
template <int NF, class S>
__device__ __forceinline__ void grid_rhs_all_f32(const GridView<S> &G, const RtThermalF &rt,
                                                 const float *s, float *ds)
{
	Stencil<float> st;
	make_stencil<float, S>(G, s[0], s[1], s[2], s[3], st);
	const float u0 = s[4], u1 = s[5], u2 = s[6], u3 = s[7];
	const float P[10] = {u0 * u0, 2.f * u0 * u1, 2.f * u0 * u2, 2.f * u0 * u3, u1 * u1,
	                     2.f * u1 * u2, 2.f * u1 * u3, u2 * u2, 2.f * u2 * u3, u3 * u3};
	float uGu[4];
	lev_contract<S>(G.gam[0], st, P, uGu);
	if (st.wt > 0.f) {   /* wt = 0: the t1 level alone (time clamped) */
		float y2[4];
		lev_contract<S>(G.gam[1], st, P, y2);
#pragma unroll
		for (int mu = 0; mu < 4; ++mu)
			uGu[mu] = uGu[mu] + st.wt * (y2[mu] - uGu[mu]);
	}
	const float u[4] = {u0, u1, u2, u3};
	ds[0] = u0; ds[1] = u1; ds[2] = u2; ds[3] = u3;
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
		ds[4 + mu] = -uGu[mu] + uGu[0] * u[mu];
	if constexpr (NF > 0) {
		float kt, ne, bkcos;
		flow_f32<S>(G, rt, st, u, kt, ne, bkcos);
		thermal_f32<NF>(rt, ne, -kt, bkcos, s + 8 + NF, ds + 8, ds + 8 + NF);
	}
}

/* Classical RK4 with stages from hi, increment added to (hi, lo). */
template <int NF, class S>
__device__ __forceinline__ void grid_rk4_f32c(const GridView<S> &G, const RtThermalF &rt,
                                              float *hi, float *lo, float dt)
{
	constexpr int ND = 8 + 2 * NF;
	float k[ND], acc[ND], x[ND];
	const float hdt = 0.5f * dt;
	grid_rhs_all_f32<NF, S>(G, rt, hi, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] = k[i];
		x[i] = hi[i] + hdt * k[i];
	}
	grid_rhs_all_f32<NF, S>(G, rt, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + hdt * k[i];
	}
	grid_rhs_all_f32<NF, S>(G, rt, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + dt * k[i];
	}
	grid_rhs_all_f32<NF, S>(G, rt, x, k);
	const float dt6 = dt * (1.0f / 6.0f);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		float y = (acc[i] + k[i]) * dt6 + lo[i];
		float t = hi[i] + y;
		lo[i] = y - (t - hi[i]);
		hi[i] = t;
	}
}

} // namespace gray
