/*
 * kernels_grid32.hip -- FP32 form of the numerical-spacetime path: one dump
 * interval per launch, FP32 RHS and RK4 stages on a compensated (hi + lo)
 * state, FP32 or FP64 grid storage; transfer channels (the reference flow)
 * on the state scaled by K as described in rt_thermal.h.
 *
 * The RHS contracts each corner's 40 Christoffels with the ten products
 * u^a u^b first and interpolates the four results (linear operations
 * commute), so a corner costs one contiguous 160 B (FP32) or 320 B (FP64)
 * read and 40 FMAs, and no per-field interpolation state is kept.
 */
#include <climits>

#include "grid_f32.h"
#include "launch.h"

namespace gray {

template <int NF, class S>
__global__ __launch_bounds__(256) void k_grid_evolve_f32(GridView<S> G, RtThermal rt64, StepRule sr,
                                                         double *__restrict__ data,
                                                         int64_t *__restrict__ steps,
                                                         size_t n_rays,
                                                         unsigned long long *__restrict__ nsteps)
{
	constexpr int ND = 8 + 2 * NF;
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	const RtThermalF rt = rt_thermal_f32(rt64);
	const double inv_scale = NF > 0 ? 1.0 / rt64.scale : 1.0;
	float hi[ND], lo[ND];
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		double v = data[(size_t)i * n_rays + g];
		if (i >= 8)
			v *= inv_scale;               /* I' = I / K, tau' = tau / K */
		hi[i] = (float)v;
		lo[i] = (float)(v - (double)hi[i]);
	}
	const float dts = (float)sr.dts;
	const double dt_over_n = sr.dt / (double)sr.n;
	int64_t nst = 0;
	int dh;
	for (int h = 0; h < sr.n; h += dh) {
		float gdt = grid_getdt<float, S>(G, hi[1], hi[2], hi[3], hi[5], hi[6], hi[7], dts);
		double xd = (double)gdt / sr.dt * (double)sr.n;
		dh = (xd == xd) ? (int)xd : 0;
		if (!dh)
			break;
		if (dh > sr.n - h)
			dh = sr.n - h;
		grid_rk4_f32c<NF, S>(G, rt, hi, lo, (float)((double)dh * dt_over_n));
		++nst;
	}
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		double v = (double)hi[i] + (double)lo[i];
		if (i >= 8)
			v *= rt64.scale;
		data[(size_t)i * n_rays + g] = v;
	}
	steps[g] += nst;
	if (nsteps)
		atomicAdd(nsteps, (unsigned long long)nst);
}

static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

template <int NF>
static void grid_evolve_f32_nf(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                               double *data, int64_t *steps, size_t n_rays,
                               unsigned long long *nsteps, hipStream_t st)
{
	if (gd.dtype == 4)
		hipLaunchKernelGGL((k_grid_evolve_f32<NF, float>), dim3(blocks_for(n_rays, 256)), dim3(256),
		                   0, st, gd.view<float>(), rt, sr, data, steps, n_rays, nsteps);
	else
		hipLaunchKernelGGL((k_grid_evolve_f32<NF, double>), dim3(blocks_for(n_rays, 256)), dim3(256),
		                   0, st, gd.view<double>(), rt, sr, data, steps, n_rays, nsteps);
}

hipError_t launch_grid_evolve_f32(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                                  double *data, int64_t *steps, size_t n_rays,
                                  unsigned long long *nsteps, hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
#define L(NF) grid_evolve_f32_nf<NF>(gd, rt, sr, data, steps, n_rays, nsteps, st)
	GRAY_NF_SWITCH(rt.n_freq, L)
#undef L
	return hipGetLastError();
}

} // namespace gray
