/*
 * kernels_fb.hip -- FP64 re-integration of the rays the FP32 pass hands
 * over (step collapse near the horizon, or a non-finite FP32 state).
 *
 * The FP64 step of step_fb.h (few instructions per step: these rays'
 * sequential steps are a render's critical path), compiled with
 * multiply-add contraction: these rays are not part of the bit-exact form.
 */
#include "launch.h"
#include "step_fb.h"

namespace gray {

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
	constexpr int ND = nd64<NF>();
	unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= ctr->n_fb)
		return;
	/* These few waves share SIMDs with the next pool pass and their
	 * longest ray is the render's critical path: raise their issue
	 * priority so the pool waves fill the gaps of their dependency chain. */
	__builtin_amdgcn_s_setprio(3);
	const RtConst64<NF> rc = make_rtconst64<NF>(p);
	Fallback f = fb[e];
	size_t g = (size_t)f.g;
	double s[ND];
	const int d0 = f.dump, h0 = f.h;
	/* stored: 8 + 2 NF reals with tau_i in place; integrated: T = tau_0 / d_0 */
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = data[(size_t)i * n_rays + g];
	if constexpr (NF > 0)
		s[8 + NF] = s[8 + NF] / rc.d[0];
	fb_lower(p.a, s);
	const int64_t st0 = steps[g];
	int64_t st = st0;
	for (int d = d0; d < sr.n_dump; ++d) {
		bool retired;
		st += interval_fb<NF>(p, rc, sr, s, d == d0 ? h0 : 0, &retired);
		if (retired)
			break;
	}
	fb_raise(p.a, s);
#pragma unroll
	for (int i = 0; i < 8 + NF; ++i)
		data[(size_t)i * n_rays + g] = s[i];
	if constexpr (NF > 0) {
#pragma unroll
		for (int i = 0; i < NF; ++i)
			data[(size_t)(8 + NF + i) * n_rays + g] = rc.d[i] * s[8 + NF];
	}
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	info[g] = getuu(p.a, q, u);
	steps[g] = st;
	atomicAdd(&ctr->fb_steps, (unsigned long long)(st - st0));
}

static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

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

} // namespace gray
