/*
 * kernels_f32.hip -- the FP32 form: a persistent ray-pool kernel.
 *
 * Each lane owns one ray at a time.  A wave refills its idle lanes together:
 * one ballot collects the lanes whose ray retired, one atomicAdd on the
 * global ray counter reserves that many ray indices, and each lane takes
 * base + (its rank among the idle lanes).  Waves exit when the counter has
 * passed the last ray and all their lanes are idle, so every wave ends.
 *
 * Arithmetic: the RK4 stages and the geodesic / transfer RHS run in FP32
 * with the hardware rcp/sqrt/exp; the state is carried as an unevaluated
 * sum hi + lo of two floats (Kahan-compensated update), which keeps the
 * position at r ~ 1000 and the slowly varying momentum accurate far below
 * one FP32 ulp per step.  Camera initial data come from the FP64 icond
 * kernel; the quantised step schedule (sim-org/driver.cl:81-99) is computed
 * in FP64.  A ray whose step falls below FB_STEP_MIN (horizon approach) is
 * handed to the FP64 kernel, which continues it from its compensated state
 * (hi + lo promoted to double); a ray whose FP32 state turns non-finite is
 * redone in FP64 from the state the pass started from.
 */
#include <climits>

#include "launch.h"
#include "step_f32.h"

namespace gray {

/* The pool's render-uniform scalars, converted on the host: each stays in
 * a scalar register, where a conversion in the kernel would hold its result
 * in a vector register for the whole kernel.  Field names are those of
 * RayParams that getdt and rhs_f32 read. */
template <int NF> struct PoolConst {
	float a, r_h, eps_stop, r_esc, r_isco;   /* (float) of RayParams' */
	int escape, fluid;
	float dts;                                 /* (float) sr.dts         */
	double n_over_dt, dt_over_n;               /* sr.n / sr.dt, sr.dt / sr.n */
	RtConst<NF> rc;
};

template <int NF>
PoolConst<NF> make_pool_const(const RayParams &p, const StepRule &sr)
{
	PoolConst<NF> c;
	c.a = (float)p.a;
	c.r_h = (float)p.r_h;
	c.eps_stop = (float)p.eps_stop;
	c.r_esc = (float)p.r_esc;
	c.r_isco = (float)p.r_isco;
	c.escape = p.escape;
	c.fluid = p.fluid;
	c.dts = (float)sr.dts;
	/* the quantised step count is gdt n / dt; the FP32 form takes it as
	 * gdt (n / dt), one multiply instead of a division per step */
	c.n_over_dt = (double)sr.n / sr.dt;
	c.dt_over_n = sr.dt / (double)sr.n;
	c.rc = make_rtconst<NF>(p);
	return c;
}

template <int NF>
__global__ __launch_bounds__(256) void k_pool_f32(RayParams p, StepRule sr,
                                                  double *__restrict__ data,
                                                  double *__restrict__ info,
                                                  int64_t *__restrict__ steps,
                                                  size_t n_rays, RayRange rr,
                                                  Counters *__restrict__ ctr,
                                                  Fallback *__restrict__ fb,
                                                  const PoolConst<NF> pc)
{
	const unsigned n0 = (unsigned)(rr.e0 - rr.b0), n_all = n0 + (unsigned)(rr.e1 - rr.b1);
	constexpr int ND = nd32<NF>();
	const unsigned lane = __lane_id();
	/* the channels' l2c in vector registers for the whole kernel (an FMA
	 * takes one scalar operand, and d2 is the other), and the spin, which
	 * the RHS reads 18 times per evaluation (an FP32 instruction reading a
	 * scalar register issues slower); the register moves are opaque to
	 * the compiler, which would otherwise repeat them inside the loop */
	RtConst<NF> rc = pc.rc;
#pragma unroll
	for (int i = 0; i < NF; ++i)
		asm("v_mov_b32 %0, %1" : "=v"(rc.l2c[i]) : "s"(pc.rc.l2c[i]));
	PoolConst<NF> pv = pc;
	asm("v_mov_b32 %0, %1" : "=v"(pv.a) : "s"(pc.a));

	float hi[ND], lo[ND];
	int g = -1;
	int d = 0, h = 0;
	int dst = 0;                 /* steps taken since the ray was loaded */
	bool exhausted = false;

	for (;;) {
		/* wave-level refill of retired lanes */
		const bool need = (g < 0) && !exhausted;
		const unsigned long long m = __ballot(need);
		if (m) {
			const int leader = __ffsll((unsigned long long)m) - 1;
			unsigned base = 0;
			if ((int)lane == leader)
				base = atomicAdd(&ctr->ray_next, (unsigned)__popcll(m));
			base = __shfl(base, leader);
			if (need) {
				unsigned idx = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
				if (idx < n_all) {
					g = idx < n0 ? rr.b0 + (int)idx : rr.b1 + (int)(idx - n0);
					load_state<NF>(p, data, n_rays, (size_t)g, hi, lo);
					dst = 0;
					d = 0;
					h = 0;
				} else {
					exhausted = true;
				}
			}
		}
		if (g < 0)
			break;

		/* the first RK4 stage: its dx/dlambda and r are the step rule's */
		float k1[ND], r1;
		rhs_f32<NF>(pv, rc, hi, k1, &r1);
		V4<float> q = {hi[0], hi[1], hi[2], hi[3]};
		V4<float> u = {k1[0], k1[1], k1[2], k1[3]};
		float gdt = getdt_r(pv, r1, q, u, pc.dts);
		double xd = (double)gdt * pc.n_over_dt;
		int dh = (xd == xd) ? (int)xd : 0;
		bool finish = (dh == 0);
		if (!finish) {
			if (dh > sr.n - h)
				dh = sr.n - h;
			double hd = (double)dh * pc.dt_over_n;
			if (fabs(hd) < FB_STEP_MIN) {
				/* step collapse: FP64 re-integration from here */
				store_state<NF>(p, data, n_rays, (size_t)g, hi, lo);
				steps[g] += dst;
				atomicAdd(&ctr->pool_steps, (unsigned long long)dst);
				unsigned slot = atomicAdd(&ctr->n_fb, 1u);
				fb[slot] = Fallback{(int)g, d, h, 0};
				g = -1;
				continue;
			}
			rk4_f32c<NF>(pv, rc, hi, lo, k1, (float)hd);
			++dst;
			float chk = ((hi[1] + hi[2]) + (hi[3] + hi[4])) + ((hi[5] + hi[6]) + hi[7]);
			if (!__builtin_isfinite(chk)) {
				/* the stored state is still the one this pass started
				 * from: redo the whole pass in FP64 */
				unsigned slot = atomicAdd(&ctr->n_fb, 1u);
				fb[slot] = Fallback{(int)g, 0, 0, 0};
				g = -1;
				continue;
			}
			h += dh;
			if (h >= sr.n) {
				h = 0;
				finish = (++d >= sr.n_dump);
			}
		}
		if (finish) {
			info[g] = store_state<NF>(p, data, n_rays, (size_t)g, hi, lo);
			steps[g] += dst;
			atomicAdd(&ctr->pool_steps, (unsigned long long)dst);
			g = -1;
		}
	}
}

/* Counters cleared by a kernel rather than a memset, so that the clear is an
 * ordinary kernel node when a caller captures the launches in a graph. */
__global__ void k_zero_counters(Counters *ctr, int n)
{
	const int i = (int)threadIdx.x;
	if (i < n)
		ctr[i] = Counters{0u, 0u, 0ull, 0ull};
}

hipError_t launch_zero_counters(Counters *ctr, int n, hipStream_t st)
{
	if (n <= 0 || n > 256)
		return hipErrorInvalidValue;
	hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(256), 0, st, ctr, n);
	return hipGetLastError();
}

hipError_t pool_f32_blocks_per_cu(int n_freq, int *blocks)
{
	*blocks = 0;
#define L(NF) return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_pool_f32<NF>, 256, 0)
	GRAY_NF_SWITCH(n_freq, L)
#undef L
	return hipSuccess;
}

hipError_t launch_pool_f32(const RayParams &p, const StepRule &sr, double *data, double *info,
                           int64_t *steps, size_t n_rays, const RayRange &rr, Counters *ctr,
                           Fallback *fb, unsigned n_blocks, hipStream_t st)
{
	if (rr.e0 - rr.b0 + rr.e1 - rr.b1 <= 0 || n_blocks == 0)
		return hipSuccess;
#define L(NF) hipLaunchKernelGGL(k_pool_f32<NF>, dim3(n_blocks), dim3(256), 0, st, \
	                         p, sr, data, info, steps, n_rays, rr, ctr, fb, make_pool_const<NF>(p, sr))
	GRAY_NF_SWITCH(p.n_freq, L)
#undef L
	return hipGetLastError();
}

} // namespace gray
