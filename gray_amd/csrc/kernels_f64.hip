/*
 * kernels_f64.hip -- FP64 kernels: camera initial conditions, the plain FP64
 * evolve (one lane per ray, the reference's per-launch structure) and the
 * FP64 re-integration of rays handed over by the FP32 pass.
 *
 * Built with -ffp-contract=off: the double arithmetic follows the written
 * order exactly (see gray_device.h).
 */
#include <climits>

#include "gray_device.h"
#include "launch.h"

namespace gray {

template <int NF>
__device__ __forceinline__ void rhs_f64(const RayParams &p, const double *s, double *ds)
{
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	V4<double> a = gr_rhs(p.a, q, u);
	ds[0] = u.t; ds[1] = u.x; ds[2] = u.y; ds[3] = u.z;
	ds[4] = a.t; ds[5] = a.x; ds[6] = a.y; ds[7] = a.z;
	if (NF > 0)
		rt_rhs<double, NF>(p, q, u, s + 8 + NF, ds + 8, ds + 8 + NF);
}

/* sim-org/RK4.cl:37-47: s + dt (k1 + 2 (k2 + k3) + k4) / 6, stages at
 * s + (dt/2) k1, s + (dt/2) k2, s + dt k3. */
template <int NF>
__device__ __forceinline__ void rk4_f64(const RayParams &p, double *s, double dt)
{
	constexpr int ND = 8 + 2 * NF;
	double k1[ND], k[ND], x[ND], k23[ND];
	const double hdt = 0.5 * dt;

	rhs_f64<NF>(p, s, k1);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = s[i] + hdt * k1[i];
	rhs_f64<NF>(p, x, k23);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = s[i] + hdt * k23[i];
	rhs_f64<NF>(p, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		k23[i] = k23[i] + k[i];
		x[i] = s[i] + dt * k[i];
	}
	rhs_f64<NF>(p, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = s[i] + (dt * ((k1[i] + 2.0 * k23[i]) + k[i])) / 6.0;
}

/* Sub-steps of one interval from unit h0 (sim-org/driver.cl:92-99).
 * Returns steps taken; *retired set when the step rule returned zero. */
template <int NF>
__device__ __forceinline__ int64_t interval_f64(const RayParams &p, const StepRule &sr,
                                                double *s, int h0, bool *retired)
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
		rk4_f64<NF>(p, s, (double)dh * sr.dt / (double)sr.n);
		++st;
	}
	return st;
}

template <int NF>
__device__ __forceinline__ void icond_ray(const RayParams &p, size_t g, double *s)
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

template <int NF>
__global__ __launch_bounds__(256) void k_icond(RayParams p, double *__restrict__ data,
                                               double *__restrict__ info,
                                               int64_t *__restrict__ steps, size_t n_rays)
{
	constexpr int ND = 8 + 2 * NF;
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	double s[ND];
	icond_ray<NF>(p, g, s);
#pragma unroll
	for (int i = 0; i < ND; ++i)
		data[(size_t)i * n_rays + g] = s[i];
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	info[g] = getuu(p.a, q, u);
	steps[g] = 0;
}

/* Plain FP64 form: one lane per ray, one interval per launch. */
template <int NF>
__global__ __launch_bounds__(256) void k_evolve_f64(RayParams p, StepRule sr,
                                                    double *__restrict__ data,
                                                    double *__restrict__ info,
                                                    int64_t *__restrict__ steps,
                                                    size_t n_rays)
{
	constexpr int ND = 8 + 2 * NF;
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	double s[ND];
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = data[(size_t)i * n_rays + g];
	bool retired;
	int64_t st = interval_f64<NF>(p, sr, s, 0, &retired);
#pragma unroll
	for (int i = 0; i < ND; ++i)
		data[(size_t)i * n_rays + g] = s[i];
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	info[g] = getuu(p.a, q, u);
	steps[g] += st;
}

/* FP64 re-integration of rays handed over by the FP32 pass: continue from
 * the stored state at unit h of interval `dump` through the remaining
 * intervals. */
template <int NF>
__global__ __launch_bounds__(256) void k_fallback_f64(RayParams p, StepRule sr,
                                                      double *__restrict__ data,
                                                      double *__restrict__ info,
                                                      int64_t *__restrict__ steps,
                                                      size_t n_rays,
                                                      const Fallback *__restrict__ fb,
                                                      Counters *__restrict__ ctr)
{
	constexpr int ND = 8 + 2 * NF;
	unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= ctr->n_fb)
		return;
	Fallback f = fb[e];
	size_t g = (size_t)f.g;
	double s[ND];
	const int d0 = f.dump, h0 = f.h;
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = data[(size_t)i * n_rays + g];
	const int64_t st0 = steps[g];
	int64_t st = st0;
	for (int d = d0; d < sr.n_dump; ++d) {
		bool retired;
		st += interval_f64<NF>(p, sr, s, d == d0 ? h0 : 0, &retired);
		if (retired)
			break;
	}
#pragma unroll
	for (int i = 0; i < ND; ++i)
		data[(size_t)i * n_rays + g] = s[i];
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	info[g] = getuu(p.a, q, u);
	steps[g] = st;
	atomicAdd(&ctr->fb_steps, (unsigned long long)(st - st0));
}

/* SoA (device) <-> AoS (dump order) in the requested real size. */
template <class R>
__global__ __launch_bounds__(256) void k_soa_to_aos(const double *__restrict__ soa,
                                                    R *__restrict__ aos, size_t n_rays, int nd)
{
	size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= n_rays * (size_t)nd)
		return;
	size_t g = idx / nd, s = idx % nd;
	aos[idx] = (R)soa[s * n_rays + g];
}

template <class R>
__global__ __launch_bounds__(256) void k_aos_to_soa(const R *__restrict__ aos,
                                                    double *__restrict__ soa, size_t n_rays, int nd)
{
	size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= n_rays * (size_t)nd)
		return;
	size_t g = idx / nd, s = idx % nd;
	soa[s * n_rays + g] = (double)aos[idx];
}

template <class R>
__global__ __launch_bounds__(256) void k_cast(const double *__restrict__ in, R *__restrict__ out, size_t n)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = (R)in[i];
}

__global__ __launch_bounds__(256) void k_info_f64(RayParams p, const double *__restrict__ data,
                                                  double *__restrict__ info, size_t n_rays)
{
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	V4<double> q = {data[g], data[n_rays + g], data[2 * n_rays + g], data[3 * n_rays + g]};
	V4<double> u = {data[4 * n_rays + g], data[5 * n_rays + g], data[6 * n_rays + g],
	                data[7 * n_rays + g]};
	info[g] = getuu(p.a, q, u);
}

static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

#define GRAY_NF_SWITCH(NFV, CALL)                       \
	switch (NFV) {                                  \
	case 0: CALL(0); break;                         \
	case 1: CALL(1); break;                         \
	case 2: CALL(2); break;                         \
	case 3: CALL(3); break;                         \
	case 4: CALL(4); break;                         \
	case 5: CALL(5); break;                         \
	case 6: CALL(6); break;                         \
	case 7: CALL(7); break;                         \
	case 8: CALL(8); break;                         \
	default: return hipErrorInvalidValue;           \
	}

hipError_t launch_icond(const RayParams &p, double *data, double *info, int64_t *steps,
                        size_t n_rays, hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
#define L(NF) hipLaunchKernelGGL(k_icond<NF>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st, \
	                         p, data, info, steps, n_rays)
	GRAY_NF_SWITCH(p.n_freq, L)
#undef L
	return hipGetLastError();
}

hipError_t launch_evolve_f64(const RayParams &p, const StepRule &sr, double *data, double *info,
                             int64_t *steps, size_t n_rays, hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
#define L(NF) hipLaunchKernelGGL(k_evolve_f64<NF>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st, \
	                         p, sr, data, info, steps, n_rays)
	GRAY_NF_SWITCH(p.n_freq, L)
#undef L
	return hipGetLastError();
}

hipError_t launch_fallback_f64(const RayParams &p, const StepRule &sr, double *data, double *info,
                               int64_t *steps, size_t n_rays, const Fallback *fb,
                               Counters *ctr, unsigned max_fb, hipStream_t st)
{
	if (max_fb == 0)
		return hipSuccess;
#define L(NF) hipLaunchKernelGGL(k_fallback_f64<NF>, dim3(blocks_for(max_fb, 256)), dim3(256), 0, st, \
	                         p, sr, data, info, steps, n_rays, fb, ctr)
	GRAY_NF_SWITCH(p.n_freq, L)
#undef L
	return hipGetLastError();
}

hipError_t launch_soa_to_aos(const double *soa, void *aos, int real_size, size_t n_rays, int nd,
                             hipStream_t st)
{
	size_t n = n_rays * (size_t)nd;
	if (n == 0)
		return hipSuccess;
	if (real_size == 8)
		hipLaunchKernelGGL(k_soa_to_aos<double>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   soa, (double *)aos, n_rays, nd);
	else
		hipLaunchKernelGGL(k_soa_to_aos<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   soa, (float *)aos, n_rays, nd);
	return hipGetLastError();
}

hipError_t launch_aos_to_soa(const void *aos, double *soa, int real_size, size_t n_rays, int nd,
                             hipStream_t st)
{
	size_t n = n_rays * (size_t)nd;
	if (n == 0)
		return hipSuccess;
	if (real_size == 8)
		hipLaunchKernelGGL(k_aos_to_soa<double>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   (const double *)aos, soa, n_rays, nd);
	else
		hipLaunchKernelGGL(k_aos_to_soa<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   (const float *)aos, soa, n_rays, nd);
	return hipGetLastError();
}

hipError_t launch_cast(const double *in, void *out, int real_size, size_t n, hipStream_t st)
{
	if (n == 0)
		return hipSuccess;
	if (real_size == 8)
		return hipMemcpyAsync(out, in, n * sizeof(double), hipMemcpyDeviceToDevice, st);
	hipLaunchKernelGGL(k_cast<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
	                   in, (float *)out, n);
	return hipGetLastError();
}

hipError_t launch_info_f64(const RayParams &p, const double *data, double *info, size_t n_rays,
                           hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_info_f64, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st,
	                   p, data, info, n_rays);
	return hipGetLastError();
}

} // namespace gray
