/*
 * kernels_grid32.hip -- FP32 form of the numerical-spacetime path: one dump
 * interval per launch, FP32 RHS and RK4 stages on a compensated (hi + lo)
 * state, FP32 or FP64 grid storage.
 *
 * The RHS contracts each corner's 40 Christoffels with the ten products
 * u^a u^b first and interpolates the four results (linear operations
 * commute), so a corner costs one contiguous 160 B (FP32) or 320 B (FP64)
 * read and 40 FMAs, and no per-field interpolation state is kept.
 */
#include <climits>

#include "grid_device.h"
#include "launch.h"

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

template <class S>
__global__ __launch_bounds__(256) void k_grid_evolve_f32(GridView<S> G, StepRule sr,
                                                         double *__restrict__ data,
                                                         int64_t *__restrict__ steps,
                                                         size_t n_rays,
                                                         unsigned long long *__restrict__ nsteps)
{
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	float hi[8], lo[8];
#pragma unroll
	for (int i = 0; i < 8; ++i) {
		double v = data[(size_t)i * n_rays + g];
		hi[i] = (float)v;
		lo[i] = (float)(v - (double)hi[i]);
	}
	const float dts = (float)sr.dts;
	const double dt_over_n = sr.dt / (double)sr.n;
	int64_t nst = 0;
	int dh;
	for (int h = 0; h < sr.n; h += dh) {
		float gdt = grid_getdt<float, S>(G, hi[1], hi[2], hi[3], dts);
		double xd = (double)gdt / sr.dt * (double)sr.n;
		dh = (xd == xd) ? (int)xd : 0;
		if (!dh)
			break;
		if (dh > sr.n - h)
			dh = sr.n - h;
		grid_rk4_f32c<S>(G, hi, lo, (float)((double)dh * dt_over_n));
		++nst;
	}
#pragma unroll
	for (int i = 0; i < 8; ++i)
		data[(size_t)i * n_rays + g] = (double)hi[i] + (double)lo[i];
	steps[g] += nst;
	if (nsteps)
		atomicAdd(nsteps, (unsigned long long)nst);
}

static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_grid_evolve_f32(const GridDesc &gd, const StepRule &sr, double *data,
                                  int64_t *steps, size_t n_rays, unsigned long long *nsteps,
                                  hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
	if (gd.dtype == 4)
		hipLaunchKernelGGL(k_grid_evolve_f32<float>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st,
		                   gd.view<float>(), sr, data, steps, n_rays, nsteps);
	else
		hipLaunchKernelGGL(k_grid_evolve_f32<double>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st,
		                   gd.view<double>(), sr, data, steps, n_rays, nsteps);
	return hipGetLastError();
}

} // namespace gray
