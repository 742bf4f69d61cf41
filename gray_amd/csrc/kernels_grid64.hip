/*
 * kernels_grid64.hip -- plain FP64 form of the numerical-spacetime path:
 * camera initial conditions with the interpolated metric, one dump interval
 * per launch (one lane per ray), and g(u,u) diagnostics.
 *
 * Built with -ffp-contract=off and evaluated in the order of the FP64 oracle
 * (interpolate each Christoffel, blend in time, then contract as
 * sim-org/dyst.cl:186-191 does), so on uniform grids it reproduces the
 * oracle bit for bit.
 */
#include <climits>

#include "grid_f64.h"
#include "launch.h"

namespace gray {

/* sim-org/RK4.cl:37-47 on the whole state (8 + 2 NF components). */
template <int NF, class S>
__device__ __forceinline__ void grid_rk4_f64(const GridView<S> &G, const RtThermal &rt, double *s,
                                             double dt)
{
	constexpr int ND = 8 + 2 * NF;
	double k1[ND], k2[ND], k[ND], x[ND];
	const double hdt = 0.5 * dt;
	grid_rhs_all_f64<NF, S>(G, rt, s, k1);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = s[i] + hdt * k1[i];
	grid_rhs_all_f64<NF, S>(G, rt, x, k2);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = s[i] + hdt * k2[i];
	grid_rhs_all_f64<NF, S>(G, rt, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		k2[i] = k2[i] + k[i];
		x[i] = s[i] + dt * k[i];
	}
	grid_rhs_all_f64<NF, S>(G, rt, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = s[i] + (dt * ((k1[i] + 2.0 * k2[i]) + k[i])) / 6.0;
}

template <class S>
__global__ __launch_bounds__(256) void k_grid_icond(GridView<S> G, RayParams p,
                                                    double *__restrict__ data,
                                                    double *__restrict__ info,
                                                    int64_t *__restrict__ steps, size_t n_rays,
                                                    int nd)
{
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	size_t gi = g % (size_t)p.w_rays, gj = g / (size_t)p.w_rays;
	double alpha = (((double)gi + 0.5) / (double)p.w_rays - 0.5) * p.w_img;
	double beta  = (((double)gj + 0.5) / (double)p.h_rays - 0.5) * p.h_img;
	double R0 = p.r_obs * p.si - beta * p.ci;
	double z  = p.r_obs * p.ci + beta * p.si;
	double y  = R0 * p.sj + alpha * p.cj;
	double x  = R0 * p.cj - alpha * p.sj;
	double s[8] = {0.0, x, y, z, 1.0, p.si * p.cj, p.si * p.sj, p.ci};
	if (!G.raw_icond) {
		Stencil<double> st;
		make_stencil<double, S>(G, s[0], s[1], s[2], s[3], st);
		double m[10];
		sample_f64<10, 10, S>(G.met, 0, st, m);
		double A = m[0];
		double B = ((m[1] * s[5] + m[2] * s[6]) + m[3] * s[7]) * 2.0;
		double C = ((((m[4] * s[5] + m[5] * s[6]) + m[6] * s[7]) * s[5]) +
		            (((m[5] * s[5] + m[7] * s[6]) + m[8] * s[7]) * s[6])) +
		            (((m[6] * s[5] + m[8] * s[6]) + m[9] * s[7]) * s[7]);
		double sc = (-(B + sqrt(B * B - (4.0 * A) * C))) / (2.0 * A);
		s[5] = s[5] / sc;
		s[6] = s[6] / sc;
		s[7] = s[7] / sc;
	}
#pragma unroll
	for (int i = 0; i < 8; ++i)
		data[(size_t)i * n_rays + g] = s[i];
	for (int i = 8; i < nd; ++i)
		data[(size_t)i * n_rays + g] = 0.0;   /* rt_icond, sim-org/rt.cl:179-183 */
	info[g] = grid_getuu_f64<S>(G, s);
	steps[g] = 0;
}

template <int NF, class S>
__global__ __launch_bounds__(256) void k_grid_evolve_f64(GridView<S> G, RtThermal rt, StepRule sr,
                                                         double *__restrict__ data,
                                                         double *__restrict__ info,
                                                         int64_t *__restrict__ steps, size_t n_rays)
{
	constexpr int ND = 8 + 2 * NF;
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	double s[ND];
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = data[(size_t)i * n_rays + g];
	int64_t nst = 0;
	int dh;
	for (int h = 0; h < sr.n; h += dh) {
		double x = grid_getdt<double, S>(G, s[1], s[2], s[3], s[5], s[6], s[7], sr.dts) / sr.dt * (double)sr.n;
		dh = (x == x) ? (int)x : 0;
		if (!dh)
			break;
		if (dh > sr.n - h)
			dh = sr.n - h;
		grid_rk4_f64<NF, S>(G, rt, s, (double)dh * sr.dt / (double)sr.n);
		++nst;
	}
#pragma unroll
	for (int i = 0; i < ND; ++i)
		data[(size_t)i * n_rays + g] = s[i];
	info[g] = grid_getuu_f64<S>(G, s);
	steps[g] += nst;
}

template <class S>
__global__ __launch_bounds__(256) void k_grid_info_f64(GridView<S> G, const double *__restrict__ data,
                                                       double *__restrict__ info, size_t n_rays)
{
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	double s[8];
#pragma unroll
	for (int i = 0; i < 8; ++i)
		s[i] = data[(size_t)i * n_rays + g];
	info[g] = grid_getuu_f64<S>(G, s);
}

/* Sample of one field at given events (test and tool helper). */
template <class S>
__global__ void k_grid_sample_f64(GridView<S> G, const double *__restrict__ ev, int which, int comp,
                                  double *__restrict__ out, int n)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	Stencil<double> st;
	make_stencil<double, S>(G, ev[4 * i], ev[4 * i + 1], ev[4 * i + 2], ev[4 * i + 3], st);
	double v[1];
	if (which == 0)
		sample_f64<1, 40, S>(G.gam, comp, st, v);
	else if (which == 1)
		sample_f64<1, 10, S>(G.met, comp, st, v);
	else
		sample_f64<1, 1, S>(G.rho, 0, st, v);
	out[i] = v[0];
}

static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_grid_icond(const GridDesc &gd, const RayParams &p, double *data, double *info,
                             int64_t *steps, size_t n_rays, hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
	const int nd = 8 + 2 * p.n_freq;
	if (gd.dtype == 4)
		hipLaunchKernelGGL(k_grid_icond<float>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st,
		                   gd.view<float>(), p, data, info, steps, n_rays, nd);
	else
		hipLaunchKernelGGL(k_grid_icond<double>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st,
		                   gd.view<double>(), p, data, info, steps, n_rays, nd);
	return hipGetLastError();
}

template <int NF>
static void grid_evolve_f64_nf(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                               double *data, double *info, int64_t *steps, size_t n_rays,
                               hipStream_t st)
{
	if (gd.dtype == 4)
		hipLaunchKernelGGL((k_grid_evolve_f64<NF, float>), dim3(blocks_for(n_rays, 256)), dim3(256),
		                   0, st, gd.view<float>(), rt, sr, data, info, steps, n_rays);
	else
		hipLaunchKernelGGL((k_grid_evolve_f64<NF, double>), dim3(blocks_for(n_rays, 256)), dim3(256),
		                   0, st, gd.view<double>(), rt, sr, data, info, steps, n_rays);
}

hipError_t launch_grid_evolve_f64(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                                  double *data, double *info, int64_t *steps, size_t n_rays,
                                  hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
#define L(NF) grid_evolve_f64_nf<NF>(gd, rt, sr, data, info, steps, n_rays, st)
	GRAY_NF_SWITCH(rt.n_freq, L)
#undef L
	return hipGetLastError();
}

hipError_t launch_grid_info_f64(const GridDesc &gd, const double *data, double *info, size_t n_rays,
                                hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
	if (gd.dtype == 4)
		hipLaunchKernelGGL(k_grid_info_f64<float>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st,
		                   gd.view<float>(), data, info, n_rays);
	else
		hipLaunchKernelGGL(k_grid_info_f64<double>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st,
		                   gd.view<double>(), data, info, n_rays);
	return hipGetLastError();
}

hipError_t launch_grid_sample_f64(const GridDesc &gd, const double *ev, int which, int comp,
                                  double *out, int n, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	if (gd.dtype == 4)
		hipLaunchKernelGGL(k_grid_sample_f64<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   gd.view<float>(), ev, which, comp, out, n);
	else
		hipLaunchKernelGGL(k_grid_sample_f64<double>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   gd.view<double>(), ev, which, comp, out, n);
	return hipGetLastError();
}

} // namespace gray
