/*
 * kernels_serve.hip -- epoch kernels of the serving pool.
 *
 * Every kernel takes a compacted list of slots, one lane per slot, and
 * advances each slot's ray by at most K RK4 steps (or to its end), so an
 * epoch is bounded and new requests get slots at the next epoch.  The FP32
 * state crosses epochs as separate hi and lo floats, which is exact: a ray
 * computes the same numbers whatever epochs, slots or neighbours it gets.
 *
 *   k_srv_init    camera initial conditions in FP64 for newly assigned slots
 *   k_srv_kerr32  analytic Kerr, FP32 form (per-lane spin, channels, camera)
 *   k_srv_kerr64  analytic Kerr, FP64 re-integration of collapsed-step rays
 *                 (the step of step_fb.h, as the render's re-integration)
 *   k_srv_grid32  gridded spacetimes, FP32 form
 *
 * Built with multiply-add contraction (these are FP32-form kernels).
 */
#include <climits>

#include "grid_f32.h"
#include "grid_f64.h"
#include "launch_serve.h"
#include "step_f32.h"
#include "step_f64.h"
#include "step_fb.h"

namespace gray {

__device__ __forceinline__ LaneKerr lane_params(const RayParams &p)
{
	LaneKerr L;
	L.a = (float)p.a;
	L.r_h = (float)p.r_h;
	L.eps_stop = (float)p.eps_stop;
	L.r_esc = (float)p.r_esc;
	L.r_isco = (float)p.r_isco;
	L.escape = p.escape;
	L.fluid = p.fluid;
	L.n_freq = p.n_freq;
#pragma unroll
	for (int i = 0; i < SRV_NF; ++i)
		L.nu[i] = (float)p.nu[i < p.n_freq ? i : 0];
	return L;
}

/* Write a finished ray into its request's output: geodesic state, the
 * request's channels (I then tau), g(u,u) and the step count. */
__device__ __forceinline__ void write_out(const SrvRequest &R, int ray, const double *s, int nf_slot,
                                          double info, int64_t steps)
{
	double *o = R.out_data + (size_t)ray * R.nd;
#pragma unroll
	for (int i = 0; i < 8; ++i)
		o[i] = s[i];
	const int nf = R.p.n_freq;
	for (int i = 0; i < nf; ++i) {
		o[8 + i] = s[8 + i];
		o[8 + nf + i] = s[8 + nf_slot + i];
	}
	R.out_info[ray] = info;
	R.out_steps[ray] = steps;
}

__device__ __forceinline__ void push(int *list, unsigned *count, int v)
{
	unsigned k = atomicAdd(count, 1u);
	list[k] = v;
}

__global__ __launch_bounds__(256) void k_srv_init(const SrvRequest *__restrict__ reqs, SrvSlots S,
                                                  const int *__restrict__ a_slot,
                                                  const int *__restrict__ a_req,
                                                  const int *__restrict__ a_ray, int n)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const int s = a_slot[i], r = a_req[i], g = a_ray[i];
	const SrvRequest &R = reqs[r];
	double st[SRV_NC];
#pragma unroll
	for (int c = 0; c < SRV_NC; ++c)
		st[c] = 0.0;
	if (R.kind == SRV_KERR) {
		icond_ray<SRV_NF, RayParams>(R.p, (size_t)g, st);
	} else {
		/* camera + null normalisation with the interpolated metric */
		const RayParams &p = R.p;
		size_t gi = (size_t)g % (size_t)p.w_rays, gj = (size_t)g / (size_t)p.w_rays;
		double alpha = (((double)gi + 0.5) / (double)p.w_rays - 0.5) * p.w_img;
		double beta  = (((double)gj + 0.5) / (double)p.h_rays - 0.5) * p.h_img;
		double R0 = p.r_obs * p.si - beta * p.ci;
		st[0] = 0.0;
		st[3] = p.r_obs * p.ci + beta * p.si;
		st[2] = R0 * p.sj + alpha * p.cj;
		st[1] = R0 * p.cj - alpha * p.sj;
		st[4] = 1.0;
		st[5] = p.si * p.cj;
		st[6] = p.si * p.sj;
		st[7] = p.ci;
		Stencil<double> sc;
		make_stencil<double, float>(R.gv, st[0], st[1], st[2], st[3], sc);
		double m[10];
		sample_f64<10, 10, float>(R.gv.met, 0, sc, m);
		double A = m[0];
		double B = ((m[1] * st[5] + m[2] * st[6]) + m[3] * st[7]) * 2.0;
		double C = ((((m[4] * st[5] + m[5] * st[6]) + m[6] * st[7]) * st[5]) +
		            (((m[5] * st[5] + m[7] * st[6]) + m[8] * st[7]) * st[6])) +
		            (((m[6] * st[5] + m[8] * st[6]) + m[9] * st[7]) * st[7]);
		double k = (-(B + sqrt(B * B - (4.0 * A) * C))) / (2.0 * A);
		st[5] /= k;
		st[6] /= k;
		st[7] /= k;
	}
#pragma unroll
	for (int c = 0; c < SRV_NC; ++c) {
		float h = (float)st[c];
		S.st[(size_t)c * S.S + s] = h;
		S.st[(size_t)(SRV_NC + c) * S.S + s] = (float)(st[c] - (double)h);
	}
	S.req[s] = r;
	S.ray[s] = g;
	S.d[s] = 0;
	S.h[s] = 0;
	S.steps[s] = 0;
}

__global__ __launch_bounds__(256) void k_srv_kerr32(const SrvRequest *__restrict__ reqs, SrvSlots S,
                                                    const int *__restrict__ list, int n, int K,
                                                    SrvLists out)
{
	int idx = blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= n)
		return;
	const int s = list[idx];
	const SrvRequest &R = reqs[S.req[s]];
	const LaneKerr P = lane_params(R.p);
	const RtConst<SRV_NF> rc = make_rtconst<SRV_NF>(P);
	const StepRule sr = R.sr;
	/* the slot holds the FP32 form's state (geodesic, intensities, the
	 * shared depth integral T in channel 0's depth place; step_f32.h) */
	constexpr int NI = nd32<SRV_NF>();
	float hi[NI], lo[NI];
#pragma unroll
	for (int c = 0; c < NI; ++c) {
		hi[c] = S.st[(size_t)c * S.S + s];
		lo[c] = S.st[(size_t)(SRV_NC + c) * S.S + s];
	}
	int d = S.d[s], h = S.h[s];
	int64_t st = S.steps[s];
	const int64_t st0 = st;
	const float dts = (float)sr.dts;
	const double dt_over_n = sr.dt / (double)sr.n;
	const double n_over_dt = (double)sr.n / sr.dt;   /* gdt n / dt as one multiply */
	int state = 0; /* 0 running, 1 finished, 2 to FP64, 3 restart in FP64 */
	for (int k = 0; k < K; ++k) {
		V4<float> q = {hi[0], hi[1], hi[2], hi[3]};
		V4<float> u = {hi[4], hi[5], hi[6], hi[7]};
		float gdt = getdt(P, q, u, dts);
		double xd = (double)gdt * n_over_dt;
		int dh = (xd == xd) ? (int)xd : 0;
		if (dh == 0) {
			state = 1;
			break;
		}
		if (dh > sr.n - h)
			dh = sr.n - h;
		double hd = (double)dh * dt_over_n;
		if (fabs(hd) < 1e-4) {
			state = 2;
			break;
		}
		rk4_f32c<SRV_NF>(P, rc, hi, lo, (float)hd);
		++st;
		float chk = ((hi[1] + hi[2]) + (hi[3] + hi[4])) + ((hi[5] + hi[6]) + hi[7]);
		if (!__builtin_isfinite(chk)) {
			state = 3;
			break;
		}
		h += dh;
		if (h >= sr.n) {
			h = 0;
			if (++d >= sr.n_dump) {
				state = 1;
				break;
			}
		}
	}
	atomicAdd(&out.ctr->steps[0], (unsigned long long)(st - st0));
	if (state == 0) {
#pragma unroll
		for (int c = 0; c < NI; ++c) {
			S.st[(size_t)c * S.S + s] = hi[c];
			S.st[(size_t)(SRV_NC + c) * S.S + s] = lo[c];
		}
		S.d[s] = d;
		S.h[s] = h;
		S.steps[s] = st;
		return;
	}
	double x[SRV_NC];
	join_state<SRV_NF>(R.p, hi, lo, x);
	if (state == 1) {
		V4<double> qd = {x[0], x[1], x[2], x[3]}, ud = {x[4], x[5], x[6], x[7]};
		write_out(R, S.ray[s], x, SRV_NF, getuu(R.p.a, qd, ud), st);
		push(out.fin, &out.ctr->n_fin, s);
		return;
	}
#pragma unroll
	for (int c = 0; c < SRV_NC; ++c)
		S.st64[(size_t)c * S.S + s] = x[c];
	S.d[s] = d;
	S.h[s] = state == 3 ? -1 : h;
	S.steps[s] = st;
	push(out.to64, &out.ctr->n_to64, s);
}

__global__ __launch_bounds__(256) void k_srv_kerr64(const SrvRequest *__restrict__ reqs, SrvSlots S,
                                                    const int *__restrict__ list, int n, int K,
                                                    SrvLists out)
{
	int idx = blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= n)
		return;
	const int s = list[idx];
	const SrvRequest &R = reqs[S.req[s]];
	const StepRule sr = R.sr;
	const RtConst64<SRV_NF> rc = make_rtconst64<SRV_NF>(R.p);
	double x[SRV_NC];
	int d = S.d[s], h = S.h[s];
	int64_t st = S.steps[s];
	if (h < 0) {
		/* non-finite FP32 state: the whole ray again in FP64 */
		icond_ray<SRV_NF, RayParams>(R.p, (size_t)S.ray[s], x);
		d = 0;
		h = 0;
		st = 0;
	} else {
#pragma unroll
		for (int c = 0; c < SRV_NC; ++c)
			x[c] = S.st64[(size_t)c * S.S + s];
	}
	/* x holds the stored layout (tau_i in place); the step integrates the
	 * channels' shared depth T = tau_0 / d_0 in tau_0's place */
	x[8 + SRV_NF] = x[8 + SRV_NF] / rc.d[0];
	const int64_t st0 = st;
	const double n_over_dt = (double)sr.n / sr.dt;
	bool fin = false;
	int dh;
	for (int k = 0; k < K; ++k) {
		V4<double> q = {x[0], x[1], x[2], x[3]};
		V4<double> u = {x[4], x[5], x[6], x[7]};
		double xd = getdt(R.p, q, u, sr.dts) * n_over_dt;
		dh = (xd == xd) ? (int)xd : 0;
		if (!dh) {
			fin = true;
			break;
		}
		if (dh > sr.n - h)
			dh = sr.n - h;
		rk4_fb<SRV_NF, RayParams>(R.p, rc, x, (double)dh * sr.dt / (double)sr.n);
		++st;
		h += dh;
		if (h >= sr.n) {
			h = 0;
			if (++d >= sr.n_dump) {
				fin = true;
				break;
			}
		}
	}
	atomicAdd(&out.ctr->steps[1], (unsigned long long)(st - st0));
	{
		const double T = x[8 + SRV_NF];
#pragma unroll
		for (int i = 0; i < SRV_NF; ++i)
			x[8 + SRV_NF + i] = rc.d[i] * T;
	}
	if (fin) {
		V4<double> qd = {x[0], x[1], x[2], x[3]}, ud = {x[4], x[5], x[6], x[7]};
		write_out(R, S.ray[s], x, SRV_NF, getuu(R.p.a, qd, ud), st);
		push(out.fin, &out.ctr->n_fin, s);
		return;
	}
#pragma unroll
	for (int c = 0; c < SRV_NC; ++c)
		S.st64[(size_t)c * S.S + s] = x[c];
	S.d[s] = d;
	S.h[s] = h;
	S.steps[s] = st;
}

__global__ __launch_bounds__(256) void k_srv_grid32(const SrvRequest *__restrict__ reqs, SrvSlots S,
                                                    const int *__restrict__ list, int n, int K,
                                                    SrvLists out)
{
	int idx = blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= n)
		return;
	const int s = list[idx];
	const SrvRequest &R = reqs[S.req[s]];
	const GridView<float> &G = R.gv;
	const StepRule sr = R.sr;
	float hi[8], lo[8];
#pragma unroll
	for (int c = 0; c < 8; ++c) {
		hi[c] = S.st[(size_t)c * S.S + s];
		lo[c] = S.st[(size_t)(SRV_NC + c) * S.S + s];
	}
	int d = S.d[s], h = S.h[s];
	int64_t st = S.steps[s];
	const int64_t st0 = st;
	const float dts = (float)sr.dts;
	const double dt_over_n = sr.dt / (double)sr.n;
	const double n_over_dt = (double)sr.n / sr.dt;   /* gdt n / dt as one multiply */
	bool fin = false;
	for (int k = 0; k < K; ++k) {
		float gdt = grid_getdt<float, float>(G, hi[1], hi[2], hi[3], hi[5], hi[6], hi[7], dts);
		double xd = (double)gdt * n_over_dt;
		int dh = (xd == xd) ? (int)xd : 0;
		if (!dh) {
			fin = true;
			break;
		}
		if (dh > sr.n - h)
			dh = sr.n - h;
		grid_rk4_f32c<0, float>(G, RtThermalF{}, hi, lo, (float)((double)dh * dt_over_n));
		++st;
		h += dh;
		if (h >= sr.n) {
			h = 0;
			if (++d >= sr.n_dump) {
				fin = true;
				break;
			}
		}
	}
	atomicAdd(&out.ctr->steps[2], (unsigned long long)(st - st0));
	if (fin) {
		double x[SRV_NC];
#pragma unroll
		for (int c = 0; c < 8; ++c)
			x[c] = (double)hi[c] + (double)lo[c];
#pragma unroll
		for (int c = 8; c < SRV_NC; ++c)
			x[c] = 0.0;
		write_out(R, S.ray[s], x, SRV_NF, grid_getuu_f64<float>(G, x), st);
		push(out.fin, &out.ctr->n_fin, s);
		return;
	}
#pragma unroll
	for (int c = 0; c < 8; ++c) {
		S.st[(size_t)c * S.S + s] = hi[c];
		S.st[(size_t)(SRV_NC + c) * S.S + s] = lo[c];
	}
	S.d[s] = d;
	S.h[s] = h;
	S.steps[s] = st;
}

static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_srv_init(const SrvRequest *reqs, const SrvSlots &S, const int *a_slot,
                           const int *a_req, const int *a_ray, int n, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_srv_init, dim3(blocks_for(n, 256)), dim3(256), 0, st, reqs, S, a_slot,
	                   a_req, a_ray, n);
	return hipGetLastError();
}

hipError_t launch_srv_epoch(int which, const SrvRequest *reqs, const SrvSlots &S, const int *list,
                            int n, int K, const SrvLists &out, hipStream_t st)
{
	if (n <= 0)
		return hipSuccess;
	if (which == 0)
		hipLaunchKernelGGL(k_srv_kerr32, dim3(blocks_for(n, 256)), dim3(256), 0, st, reqs, S, list,
		                   n, K, out);
	else if (which == 1)
		hipLaunchKernelGGL(k_srv_kerr64, dim3(blocks_for(n, 256)), dim3(256), 0, st, reqs, S, list,
		                   n, K, out);
	else
		hipLaunchKernelGGL(k_srv_grid32, dim3(blocks_for(n, 256)), dim3(256), 0, st, reqs, S, list,
		                   n, K, out);
	return hipGetLastError();
}

} // namespace gray
