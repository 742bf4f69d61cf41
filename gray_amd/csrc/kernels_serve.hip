/*
 * kernels_serve.hip -- the serving pool's epoch kernels (serve.h).
 *
 *   k_srv_kerr   analytic Kerr, FP32 form: a persistent ray pool whose
 *                blocks are shared out among the Kerr classes; a wave
 *                refills its idle lanes from its class's queue with one
 *                ballot and one atomicAdd (resumed rays first, then fresh
 *                rays, whose camera initial conditions it computes in FP64);
 *                collapsed-step and non-finite rays go to the FP64 list
 *   k_srv_fb64   FP64 re-integration of that list (the step of step_fb.h,
 *                as the render's re-integration); survivors are appended to
 *                the next epoch's list, beside the Kerr kernel's hand-overs
 *   k_srv_grid   gridded spacetimes, FP32 form, one 8x8-ray tile per wave
 *                with the wave's LDS brick (grid_brick32.h), transfer
 *                channels of the reference flow on the scaled state
 *
 * Every wave runs at most K iterations, then saves what it still holds to
 * the resume lists, so an epoch is bounded and new requests enter at the
 * next one.  Saved states are the FP32 (hi, lo) pairs and integers as they
 * were: a ray computes the same numbers whatever epochs, lanes or
 * neighbours it gets.  Built with multiply-add contraction (FP32-form code).
 */
#include <climits>

#include "grid_brick32.h"
#include "grid_f64.h"
#include "launch_serve.h"
#include "step_f32.h"
#include "step_f64.h"
#include "step_fb.h"

namespace gray {

/* What the FP32 Kerr step reads (getdt, rt_rhs_f32): spin-derived values
 * per lane, the class's escape rule and fluid per block. */
struct LaneKerr {
	float a, r_h, eps_stop, r_isco, r_esc;
	int escape, fluid;
};

/* A finished ray into its request's output: geodesic state, the request's
 * channels (I, then tau: x holds I at 8 and tau at 8 + nf_slot), g(u,u), the
 * step count; the request is reported when its last ray is in. */
__device__ __forceinline__ void finish_ray(const SrvRequest &R, int tab, int ray, const double *x,
                                           int nf_slot, double info, int64_t steps,
                                           unsigned *remaining, SrvCtr *ctr, int *done)
{
	double *o = R.out_data + (size_t)ray * (size_t)R.nd;
#pragma unroll
	for (int i = 0; i < 8; ++i)
		o[i] = x[i];
	const int nf = R.p.n_freq;
	for (int i = 0; i < nf; ++i) {
		o[8 + i] = x[8 + i];
		o[8 + nf + i] = x[8 + nf_slot + i];
	}
	R.out_info[ray] = info;
	R.out_steps[ray] = steps;
	if (atomicSub(&remaining[tab], 1u) == 1u)
		done[atomicAdd(&ctr->n_done, 1u)] = tab;
}

/* Queue item f of the fresh part: segment j with start[j] <= f <
 * start[j + 1] in [s0, s1) (binary search), item first[j] + f - start[j]. */
__device__ __forceinline__ void seg_item(const SrvSegs &S, int s0, int s1, unsigned f, int &tab,
                                         int &first)
{
	int lo = s0, hi = s1 - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (S.start[mid] <= f)
			lo = mid;
		else
			hi = mid - 1;
	}
	tab = S.tab[lo];
	first = S.first[lo] + (int)(f - S.start[lo]);
}

/* The epoch's stop time, as the clear kernel set it before this launch;
 * a wave that starts at or after it is counted in late[kind] (the host
 * learns how much of a launch could not run beside the others). */
__device__ __forceinline__ unsigned long long epoch_end(SrvCtr *ctr, int kind)
{
	const unsigned long long e = *(const volatile unsigned long long *)&ctr->t_end;
	const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)e);
	const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(e >> 32));
	const unsigned long long end = ((unsigned long long)hi << 32) | lo;
	if (__lane_id() == 0 && (unsigned long long)wall_clock64() >= end)
		atomicAdd(&ctr->late[kind], 1u);
	return end;
}

/* ---------------- analytic Kerr, FP32 form ---------------- */

/* Wave-wide claim of one item per asking lane from a cursor: the lane's
 * cursor value (UINT_MAX for a lane not asking). */
__device__ __forceinline__ unsigned claim_items(unsigned *cursor, bool want)
{
	const unsigned long long m = __ballot(want);
	if (!m)
		return UINT_MAX;
	const unsigned lane = __lane_id();
	const int leader = __ffsll((long long)m) - 1;
	unsigned base = 0;
	if ((int)lane == leader)
		base = atomicAdd(cursor, (unsigned)__popcll(m));
	base = __shfl(base, leader);
	return want ? base + (unsigned)__popcll(m & ((1ull << lane) - 1ull)) : UINT_MAX;
}

/* Slot of the next entry written into region [base, base + cap) of a list,
 * from the back; -1 (counted as an overflow) past its start. */
__device__ __forceinline__ int region_slot(unsigned *counter, int base, int cap, SrvCtr *ctr)
{
	const unsigned k = atomicAdd(counter, 1u);
	if (k >= (unsigned)cap) {
		atomicAdd(&ctr->overflow, 1u);
		return -1;
	}
	return base + cap - 1 - (int)k;
}

/* Kerr resume entry i of list `in` into slot o of list `out`, unchanged. */
__device__ __forceinline__ void copy_resume(const SrvResume &in, int i, const SrvResume &out, int o)
{
	out.meta[o] = in.meta[i];
	out.steps[o] = in.steps[i];
#pragma unroll
	for (int r = 0; r < 2 * SRV_NK; ++r)
		out.st[(size_t)r * out.cap + o] = in.st[(size_t)r * in.cap + i];
}

template <int NF>
__device__ __forceinline__ void kerr_lanes(const SrvKerrArgs &A,
                                           const SrvRequest *__restrict__ reqs,
                                           const SrvClass &C, int c)
{
	constexpr int ND = nd32<NF>();
	/* the class's constants, converted on the host (scalar registers);
	 * l2c in vector registers, since the channel FMA's other constant
	 * operand is d2 */
	RtConst<NF> rc;
#pragma unroll
	for (int i = 0; i < NF; ++i) {
		asm("v_mov_b32 %0, %1" : "=v"(rc.l2c[i]) : "s"(C.l2c[i]));
		rc.d2[i] = C.d2[i];
	}
	const StepRule sr = C.sr;
	const float dts = C.dts;
	const double dt_over_n = C.dt_over_n;
	const double n_over_dt = C.n_over_dt;   /* gdt n / dt as one multiply */
	const unsigned n_resume = (unsigned)C.n_resume, n_fresh = C.n_fresh;
	LaneKerr P;
	P.r_esc = (float)C.r_esc;
	P.escape = C.escape;
	P.fluid = C.fluid;
	P.a = P.r_h = P.eps_stop = P.r_isco = 0.f;
	float hi[ND], lo[ND];
	int tab = -1, ray = 0, d = 0, h = 0;
	int64_t st = 0, st0 = 0;
	bool exhausted = false;
	bool res_left = n_resume > 0;   /* uniform: resumed rays may be unclaimed */
	const unsigned long long end = epoch_end(A.ctr, 0);
	for (int it = 0;; ++it) {
		const bool stop = it >= A.K || (unsigned long long)wall_clock64() >= end;
		bool need = tab < 0 && !exhausted && !stop;
		const bool took = need;   /* a lane asking gets a ray or is exhausted */
		/* refill: resumed rays first, then fresh ones */
		if (res_left) {
			const unsigned idx = claim_items(&A.ctr->res_claimed[c], need);
			if (__ballot(need && idx >= n_resume))
				res_left = false;
			if (need && idx < n_resume) {
				const int e = C.res_off + (int)idx;
				const int4 mt = A.in.meta[e];
				tab = mt.x;
				ray = mt.y;
				d = mt.z;
				h = mt.w;
				st = A.in.steps[e];
#pragma unroll
				for (int i = 0; i < ND; ++i) {
					hi[i] = A.in.st[(size_t)i * A.in.cap + e];
					lo[i] = A.in.st[(size_t)(SRV_NK + i) * A.in.cap + e];
				}
				need = false;
			}
		}
		const unsigned fidx = claim_items(&A.ctr->fresh_claimed[c], need);
		if (need) {
			if (fidx >= n_fresh) {
				exhausted = true;
			} else {
				seg_item(A.seg, C.seg0, C.seg1, fidx, tab, ray);
				double x[8 + 2 * NF];
				icond_ray<NF, RayParams>(reqs[tab].p, (size_t)ray, x);
				split_state<NF, SrvClass>(C, reqs[tab].p.a, x, hi, lo);
				d = 0;
				h = 0;
				st = 0;
			}
		}
		if (tab < 0)
			break;
		if (took) {
			const RayParams &p = reqs[tab].p;
			P.a = (float)p.a;
			P.r_h = (float)p.r_h;
			P.eps_stop = (float)p.eps_stop;
			P.r_isco = (float)p.r_isco;
			st0 = st;
		}
		if (stop) {
			/* the ray goes on next epoch */
			const int e = region_slot(&A.ctr->saved[c], C.out_base, C.out_cap, A.ctr);
			if (e >= 0) {
				A.out.meta[e] = make_int4(tab, ray, d, h);
				A.out.steps[e] = st;
#pragma unroll
				for (int i = 0; i < ND; ++i) {
					A.out.st[(size_t)i * A.out.cap + e] = hi[i];
					A.out.st[(size_t)(SRV_NK + i) * A.out.cap + e] = lo[i];
				}
			}
			atomicAdd(&A.ctr->steps[0], (unsigned long long)(st - st0));
			break;
		}
		/* the first RK4 stage: its dx/dlambda and r are the step rule's */
		float k1[ND], r1;
		rhs_f32<NF>(P, rc, hi, k1, &r1);
		V4<float> q = {hi[0], hi[1], hi[2], hi[3]};
		V4<float> u = {k1[0], k1[1], k1[2], k1[3]};
		const float gdt = getdt_r(P, r1, q, u, dts);
		const double xd = (double)gdt * n_over_dt;
		int dh = (xd == xd) ? (int)xd : 0;
		int state = 0;   /* 1 finished, 2 to FP64 here, 3 FP64 from the camera */
		if (dh == 0) {
			state = 1;
		} else {
			if (dh > sr.n - h)
				dh = sr.n - h;
			const double hd = (double)dh * dt_over_n;
			if (fabs(hd) < FB_STEP_MIN) {
				state = 2;
			} else {
				rk4_f32c<NF>(P, rc, hi, lo, k1, (float)hd);
				++st;
				const float chk = ((hi[1] + hi[2]) + (hi[3] + hi[4])) + ((hi[5] + hi[6]) + hi[7]);
				if (!__builtin_isfinite(chk)) {
					state = 3;
				} else {
					h += dh;
					if (h >= sr.n) {
						h = 0;
						if (++d >= sr.n_dump)
							state = 1;
					}
				}
			}
		}
		if (state == 0)
			continue;
		atomicAdd(&A.ctr->steps[0], (unsigned long long)(st - st0));
		const SrvRequest &R = reqs[tab];
		if (state == 1) {
			double x[8 + 2 * NF];
			join_state<NF, SrvClass>(C, R.p.a, hi, lo, x);
			V4<double> qd = {x[0], x[1], x[2], x[3]}, ud = {x[4], x[5], x[6], x[7]};
			finish_ray(R, tab, ray, x, NF, getuu(R.p.a, qd, ud), st, A.remaining, A.ctr, A.done);
		} else {
			/* the FP64 list's layout: t, x, y, z, p_mu (the Hamiltonian
			 * state, as here), SRV_NF intensities (unused ones 0), the shared
			 * depth T itself (converted to the channels' depths only when
			 * the ray ends, so that epochs in FP64 round nothing) */
			double x[SRV_NC];
#pragma unroll
			for (int i = 0; i < 8 + NF; ++i)
				x[i] = (double)hi[i] + (double)lo[i];
#pragma unroll
			for (int i = 8 + NF; i < SRV_NC; ++i)
				x[i] = 0.0;
			x[8 + SRV_NF] = NF > 0 ? (double)hi[ND - 1] + (double)lo[ND - 1] : 0.0;
			/* the host plans the list's room; a count past it is reported
			 * as an error and nothing is written out of bounds */
			const int e = region_slot(&A.ctr->fb_saved, 0, A.fb.cap, A.ctr);
			if (e >= 0) {
				A.fb.meta[e] = make_int4(tab, ray, d, state == 3 ? -1 : h);
				A.fb.steps[e] = st;
#pragma unroll
				for (int i = 0; i < SRV_NC; ++i)
					A.fb.x[(size_t)i * A.fb.cap + e] = x[i];
			}
		}
		tab = -1;
	}
	/* carry the resumed rays nobody claimed over to the next epoch */
	while (res_left) {
		const unsigned idx = claim_items(&A.ctr->res_claimed[c], true);
		if (__ballot(idx >= n_resume))
			res_left = false;
		if (idx < n_resume) {
			const int o = region_slot(&A.ctr->saved[c], C.out_base, C.out_cap, A.ctr);
			if (o >= 0)
				copy_resume(A.in, C.res_off + (int)idx, A.out, o);
		}
	}
}


/* The request and class tables come as restrict parameters besides A: the
 * kernels write only through other pointers, so their loads stay scalar. */
__global__ __launch_bounds__(SRV_BLOCK) void k_srv_kerr(SrvKerrArgs A,
                                                        const SrvRequest *__restrict__ reqs,
                                                        const SrvClass *__restrict__ cls)
{
	const int b = (int)blockIdx.x;
	for (int c = 0; c < A.n_cls; ++c) {
		const SrvClass &C = cls[c];
		if (b < C.block0 || b >= C.block1)
			continue;
		/* two step widths: up to 2 channels and up to 4 (the extra
		 * channels repeat channel 0 and are not reported; a class without
		 * channels has fluid none, so its channels stay zero).  With three
		 * or more instantiations inlined here the kernel needs 170
		 * registers instead of 159-161: two waves per SIMD instead of
		 * three */
		if (C.n_freq <= 2)
			kerr_lanes<2>(A, reqs, C, c);
		else
			kerr_lanes<4>(A, reqs, C, c);
		return;
	}
}

/* ---------------- FP64 re-integration ---------------- */

/* The last FP64 or grid block to start tells the host, which then launches
 * the Kerr kernel into the room left (host memory, written once). */
__device__ __forceinline__ void heavy_started(const SrvHeavy &H, SrvCtr *ctr)
{
	if (threadIdx.x == 0 && atomicAdd(&ctr->heavy_started, 1u) + 1u == H.total) {
		*H.started = H.epoch;
		__threadfence_system();
	}
}

/* FP64 entry i of list `in` into slot o of list `out`, unchanged. */
__device__ __forceinline__ void copy_fb(const SrvFb &in, int i, const SrvFb &out, int o)
{
	out.meta[o] = in.meta[i];
	out.steps[o] = in.steps[i];
#pragma unroll
	for (int c = 0; c < SRV_NC; ++c)
		out.x[(size_t)c * out.cap + o] = in.x[(size_t)c * in.cap + i];
}

/* FP64 re-integration: a persistent pool whose lanes refill from the list
 * as rays finish; at the stop time survivors, then the entries nobody
 * claimed, go to the out-list. */
__global__ __launch_bounds__(SRV_BLOCK) void k_srv_fb64(SrvFbArgs A,
                                                        const SrvRequest *__restrict__ reqs)
{
	heavy_started(A.heavy, A.ctr);
	const unsigned long long end = epoch_end(A.ctr, 1);
	const unsigned n_in = (unsigned)A.n_in;
	int tab = -1, ray = 0, d = 0, h = 0;
	double x[SRV_NC];
	int64_t st = 0, st0 = 0;
	bool exhausted = false;
	for (int it = 0;; ++it) {
		const bool stop = it >= A.K || (unsigned long long)wall_clock64() >= end;
		const bool need = tab < 0 && !exhausted && !stop;
		const unsigned idx = claim_items(&A.ctr->fb_claimed, need);
		if (need) {
			if (idx >= n_in) {
				exhausted = true;
			} else {
				const int e = A.in_off + (int)idx;
				const int4 mt = A.in.meta[e];
				tab = mt.x;
				ray = mt.y;
				d = mt.z;
				h = mt.w;
				st = A.in.steps[e];
				if (h < 0) {
					/* non-finite FP32 state: the whole ray again in FP64 */
					icond_ray<SRV_NF, RayParams>(reqs[tab].p, (size_t)ray, x);
					fb_lower(reqs[tab].p.a, x);
					d = 0;
					h = 0;
					st = 0;
				} else {
#pragma unroll
					for (int c = 0; c < SRV_NC; ++c)
						x[c] = A.in.x[(size_t)c * A.in.cap + e];
				}
				st0 = st;
			}
		}
		if (tab < 0)
			break;
		const SrvRequest &R = reqs[tab];
		if (stop) {
			const int o = region_slot(&A.ctr->fb_saved, 0, A.out.cap, A.ctr);
			if (o >= 0) {
				A.out.meta[o] = make_int4(tab, ray, d, h);
				A.out.steps[o] = st;
#pragma unroll
				for (int c = 0; c < SRV_NC; ++c)
					A.out.x[(size_t)c * A.out.cap + o] = x[c];
			}
			atomicAdd(&A.ctr->steps[1], (unsigned long long)(st - st0));
			break;
		}
		/* x holds t, x, y, z, p_mu, the intensities and at 8 + SRV_NF the
		 * channels' shared depth T (0 from the camera) */
		const StepRule sr = R.sr;
		const RtConst64<SRV_NF> rc = make_rtconst64<SRV_NF>(R.p);
		const int dh = step_fb<SRV_NF, RayParams>(R.p, rc, sr, x, h);
		bool fin = !dh;
		if (!fin) {
			++st;
			h += dh;
			if (h >= sr.n) {
				h = 0;
				fin = ++d >= sr.n_dump;
			}
		}
		if (fin) {
			atomicAdd(&A.ctr->steps[1], (unsigned long long)(st - st0));
			fb_raise(R.p.a, x);
			const double T = x[8 + SRV_NF];
#pragma unroll
			for (int i = 0; i < SRV_NF; ++i)
				x[8 + SRV_NF + i] = rc.d[i] * T;
			V4<double> qd = {x[0], x[1], x[2], x[3]}, ud = {x[4], x[5], x[6], x[7]};
			finish_ray(R, tab, ray, x, SRV_NF, getuu(R.p.a, qd, ud), st, A.remaining, A.ctr,
			           A.done);
			tab = -1;
		}
	}
	/* carry the entries nobody claimed over to the next epoch */
	bool left = !exhausted;
	while (__ballot(left)) {
		const unsigned idx = claim_items(&A.ctr->fb_claimed, true);
		if (__ballot(idx >= n_in))
			left = false;
		if (idx < n_in) {
			const int o = region_slot(&A.ctr->fb_saved, 0, A.out.cap, A.ctr);
			if (o >= 0)
				copy_fb(A.in, A.in_off + (int)idx, A.out, o);
		}
	}
}

/* ---------------- gridded spacetimes, FP32 form ---------------- */

/* Camera and null normalisation with the interpolated metric in FP64 (the
 * render's k_grid_icond). */
__device__ __forceinline__ void grid_icond_ray(const SrvRequest &R, int g, double *s)
{
	const RayParams &p = R.p;
	const int gi = g % p.w_rays, gj = g / p.w_rays;
	const double alpha = (((double)gi + 0.5) / (double)p.w_rays - 0.5) * p.w_img;
	const double beta = (((double)gj + 0.5) / (double)p.h_rays - 0.5) * p.h_img;
	const double R0 = p.r_obs * p.si - beta * p.ci;
	s[0] = 0.0;
	s[3] = p.r_obs * p.ci + beta * p.si;
	s[2] = R0 * p.sj + alpha * p.cj;
	s[1] = R0 * p.cj - alpha * p.sj;
	s[4] = 1.0;
	s[5] = p.si * p.cj;
	s[6] = p.si * p.sj;
	s[7] = p.ci;
	if (!R.gv.raw_icond) {
		Stencil<double> sc;
		make_stencil<double, float>(R.gv, s[0], s[1], s[2], s[3], sc);
		double m[10];
		sample_f64<10, 10, float>(R.gv.met, 0, sc, m);
		const double A = m[0];
		const double B = ((m[1] * s[5] + m[2] * s[6]) + m[3] * s[7]) * 2.0;
		const double Cc = ((((m[4] * s[5] + m[5] * s[6]) + m[6] * s[7]) * s[5]) +
		                   (((m[5] * s[5] + m[7] * s[6]) + m[8] * s[7]) * s[6])) +
		                  (((m[6] * s[5] + m[8] * s[6]) + m[9] * s[7]) * s[7]);
		const double k = (-(B + sqrt(B * B - (4.0 * A) * Cc))) / (2.0 * A);
		s[5] = s[5] / k;
		s[6] = s[6] / k;
		s[7] = s[7] / k;
	}
}

/* One tile of a gridded request from its current state until its rays are
 * done, the wave's iterations (it counts them) run out or the launch's stop
 * time passes; then saves the live lanes.  Returns whether it saved. */
template <int NF>
__device__ __forceinline__ bool grid_tile(const SrvGridArgs &A, const SrvRequest *__restrict__ reqs,
                                          int tab, int tile, bool resumed, int rec, float4 *brick,
                                          BrickHdr *hdr, int &it, unsigned long long end, int nf)
{
	constexpr int ND = 8 + 2 * NF;
	const unsigned lane = __lane_id();
	const SrvRequest &R = reqs[tab];
	/* the spacetime read through the table (uniform loads): a copy would
	 * hold ~70 more registers and halve the kernel's occupancy */
	const GridView<float> &G = R.gv;
	const StepRule sr = R.sr;
	const RtThermalF rt = rt_thermal_f32(R.rt);
	const double scale = R.rt.scale;
	float hi[ND > 0 ? ND : 1], lo[ND > 0 ? ND : 1];
	int ray = -1, d = 0, h = 0;
	int64_t st = 0;
	if (resumed) {
		const int e = rec * 64 + (int)lane;
		const int4 mt = A.in.meta[e];
		ray = mt.x;
		d = mt.y;
		h = mt.z;
		st = A.in.steps[e];
		if (ray >= 0) {
#pragma unroll
			for (int i = 0; i < ND; ++i) {
				hi[i] = A.in.st[(size_t)i * (64 * (size_t)A.in.cap) + e];
				lo[i] = A.in.st[(size_t)(SRV_NG + i) * (64 * (size_t)A.in.cap) + e];
			}
		}
	} else {
		const int tx = tile % R.tw, ty = tile / R.tw;
		const int x = tx * 8 + (int)(lane & 7u), y = ty * 8 + (int)(lane >> 3);
		if (x < R.p.w_rays && y < R.p.h_rays) {
			ray = y * R.p.w_rays + x;
			double s[8];
			grid_icond_ray(R, ray, s);
#pragma unroll
			for (int i = 0; i < ND; ++i) {
				const double v = i < 8 ? s[i] : 0.0;
				hi[i] = (float)v;
				lo[i] = (float)(v - (double)hi[i]);
			}
		}
	}
	/* a fresh tile starts with an empty brick (its first evaluation loads);
	 * a resumed one with the brick it had when it was saved */
	const bool use_lds = G.n[0] >= GB && G.n[1] >= GB && G.n[2] >= GB;
	int4 bo = make_int4(-1000000, -1000000, -1000000, 0);
	if (resumed)
		bo = A.in.brick[rec];
	if (lane == 0) {
		hdr->o[0] = bo.x;
		hdr->o[1] = bo.y;
		hdr->o[2] = bo.z;
		hdr->cool = bo.w;
	}
	if (use_lds && bo.x >= 0) {
		const uint32_t nx = (uint32_t)G.n[0], nxy = nx * (uint32_t)G.n[1];
		for (int q = (int)lane; q < GP_PTS; q += 64) {
			const uint32_t pt = (uint32_t)(bo.z + q / (GB * GB)) * nxy +
			                    (uint32_t)(bo.y + (q / GB) % GB) * nx + (uint32_t)(bo.x + q % GB);
			brick_store<float>(brick, q, G.gam[0] + (size_t)pt * 40);
		}
	}
	wave_sync();
	const float dts = (float)sr.dts;
	const double dt_over_n = sr.dt / (double)sr.n;
	const double n_over_dt = (double)sr.n / sr.dt;
	const int64_t st0 = st;
	bool live = ray >= 0;
	for (; it < A.K; ++it) {
		if (!__ballot(live) || (unsigned long long)wall_clock64() >= end)
			break;
		if (live) {
			const float gdt = grid_getdt<float, float>(G, hi[1], hi[2], hi[3], hi[5], hi[6],
			                                           hi[7], dts);
			const double xd = (double)gdt * n_over_dt;
			int dh = (xd == xd) ? (int)xd : 0;
			bool fin = !dh;
			if (!fin) {
				if (dh > sr.n - h)
					dh = sr.n - h;
				grid_rk4_brick<NF, float>(G, rt, brick, hdr, use_lds, hi, lo,
				                          (float)((double)dh * dt_over_n), nf);
				++st;
				h += dh;
				if (h >= sr.n) {
					h = 0;
					fin = ++d >= sr.n_dump;
				}
			}
			if (fin) {
				double x[ND];
#pragma unroll
				for (int i = 0; i < ND; ++i) {
					double v = (double)hi[i] + (double)lo[i];
					x[i] = i >= 8 ? v * scale : v;
				}
				finish_ray(R, tab, ray, x, NF, grid_getuu_f64<float>(G, x), st, A.remaining,
				           A.ctr, A.done);
				live = false;
				ray = -1;
			}
		}
	}
	atomicAdd(&A.ctr->steps[2], (unsigned long long)(st - st0));
	wave_sync();   /* the brick header as the last evaluation left it */
	const bool save = __ballot(live) != 0;
	if (save) {
		/* iterations or time used up: the tile goes on next epoch */
		int r = 0;
		if (lane == 0)
			r = region_slot(&A.ctr->g_saved, 0, A.out_cap, A.ctr);
		r = __builtin_amdgcn_readfirstlane(__shfl(r, 0));
		if (r < 0)
			return true;   /* counted as an overflow */
		const int e = r * 64 + (int)lane;
		if (lane == 0) {
			A.out.hdr[r] = make_int2(tab, tile);
			A.out.brick[r] = make_int4(hdr->o[0], hdr->o[1], hdr->o[2], hdr->cool);
		}
		A.out.meta[e] = make_int4(live ? ray : -1, d, h, 0);
		A.out.steps[e] = st;
		if (live) {
#pragma unroll
			for (int i = 0; i < ND; ++i) {
				A.out.st[(size_t)i * (64 * (size_t)A.out.cap) + e] = hi[i];
				A.out.st[(size_t)(SRV_NG + i) * (64 * (size_t)A.out.cap) + e] = lo[i];
			}
		}
	}
	return save;
}

/* Grid record i of list `in` into record o of list `out`, unchanged (the
 * whole wave copies its 64 lanes). */
__device__ __forceinline__ void copy_tile(const SrvTiles &in, int i, const SrvTiles &out, int o)
{
	const unsigned lane = __lane_id();
	if (lane == 0) {
		out.hdr[o] = in.hdr[i];
		out.brick[o] = in.brick[i];
	}
	const int a = i * 64 + (int)lane, b = o * 64 + (int)lane;
	out.meta[b] = in.meta[a];
	out.steps[b] = in.steps[a];
#pragma unroll
	for (int r = 0; r < 2 * SRV_NG; ++r)
		out.st[(size_t)r * (64 * (size_t)out.cap) + b] = in.st[(size_t)r * (64 * (size_t)in.cap) + a];
}

/* Every gridded request's tiles, built for SRV_NF channels and run with the
 * request's count (uniform per wave; the others cost no transcendentals):
 * two waves per SIMD, as the LDS bricks allow.  A wave claims tiles
 * (resumed records first, then fresh tiles) until none are left, its
 * current tile is saved or the stop time passes; then it carries the
 * resumed records nobody claimed over. */
__global__ __launch_bounds__(64 * SRV_GWAVES) __attribute__((amdgpu_waves_per_eu(2)))
void k_srv_grid(SrvGridArgs A, const SrvRequest *__restrict__ reqs)
{
	__shared__ float4 s_brick[SRV_GWAVES][GP_PTS * 10];
	__shared__ BrickHdr s_hdr[SRV_GWAVES];
	heavy_started(A.heavy, A.ctr);
	const unsigned lane = __lane_id();
	const int wv = (int)(threadIdx.x >> 6);
	float4 *brick = s_brick[wv];
	BrickHdr *hdr = &s_hdr[wv];
	int it = 0;
	const unsigned n_resume = (unsigned)A.n_resume;
	bool res_left = n_resume > 0;
	const unsigned long long end = epoch_end(A.ctr, 2);
	while (it < A.K && (unsigned long long)wall_clock64() < end) {
		int tab = -1, tile = 0, rec = -1;
		unsigned r = UINT_MAX;
		if (res_left) {
			if (lane == 0)
				r = atomicAdd(&A.ctr->g_res, 1u);
			r = __builtin_amdgcn_readfirstlane(__shfl(r, 0));
			if (r < n_resume) {
				rec = A.res_off + (int)r;
				const int2 hd = A.in.hdr[rec];
				tab = hd.x;
				tile = hd.y;
			} else {
				res_left = false;
			}
		}
		if (rec < 0) {
			unsigned f = 0;
			if (lane == 0)
				f = atomicAdd(&A.ctr->g_fresh, 1u);
			f = __builtin_amdgcn_readfirstlane(__shfl(f, 0));
			if (f >= A.n_fresh)
				break;
			seg_item(A.seg, A.seg0, A.seg1, f, tab, tile);
		}
		tab = __builtin_amdgcn_readfirstlane(tab);
		tile = __builtin_amdgcn_readfirstlane(tile);
		const int nf = __builtin_amdgcn_readfirstlane(reqs[tab].p.n_freq);
		if (grid_tile<SRV_NF>(A, reqs, tab, tile, rec >= 0, rec, brick, hdr, it, end, nf))
			break;
	}
	/* carry the resumed records nobody claimed over to the next epoch */
	while (res_left) {
		unsigned r = 0;
		if (lane == 0)
			r = atomicAdd(&A.ctr->g_res, 1u);
		r = __builtin_amdgcn_readfirstlane(__shfl(r, 0));
		if (r >= n_resume)
			break;
		int o = 0;
		if (lane == 0)
			o = region_slot(&A.ctr->g_saved, 0, A.out_cap, A.ctr);
		o = __builtin_amdgcn_readfirstlane(__shfl(o, 0));
		if (o >= 0)
			copy_tile(A.in, A.res_off + (int)r, A.out, o);
	}
}

/* Clears the epoch's counters and starts its clock: the Kerr and grid
 * launches stop at t_end, so that every wave stops at about the same
 * moment (a count of iterations would leave the fast waves idle while the
 * slowest finishes, and would run blocks that start late past the others). */
__global__ void k_srv_clear(SrvCtr *ctr, unsigned long long budget)
{
	if (threadIdx.x == 0) {
		SrvCtr c{};
		c.t_end = budget ? (unsigned long long)wall_clock64() + budget : ~0ull;
		*ctr = c;
	}
}

/* ---------------- launchers ---------------- */

hipError_t launch_srv_clear(SrvCtr *ctr, unsigned long long budget, hipStream_t st)
{
	hipLaunchKernelGGL(k_srv_clear, dim3(1), dim3(64), 0, st, ctr, budget);
	return hipGetLastError();
}

hipError_t launch_srv_kerr(const SrvKerrArgs &A, int n_blocks, hipStream_t st)
{
	if (n_blocks <= 0 || A.n_cls <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_srv_kerr, dim3(n_blocks), dim3(SRV_BLOCK), 0, st, A, A.reqs, A.cls);
	return hipGetLastError();
}

hipError_t launch_srv_fb64(const SrvFbArgs &A, int n_blocks, hipStream_t st)
{
	if (n_blocks <= 0 || A.n_in <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_srv_fb64, dim3(n_blocks), dim3(SRV_BLOCK), 0, st, A, A.reqs);
	return hipGetLastError();
}

hipError_t launch_srv_grid(const SrvGridArgs &A, int n_blocks, hipStream_t st)
{
	if (n_blocks <= 0 || (A.n_resume <= 0 && A.n_fresh == 0))
		return hipSuccess;
	hipLaunchKernelGGL(k_srv_grid, dim3(n_blocks), dim3(64 * SRV_GWAVES), 0, st, A, A.reqs);
	return hipGetLastError();
}

hipError_t srv_kerr_blocks_per_cu(int *blocks)
{
	return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_srv_kerr, SRV_BLOCK, 0);
}

hipError_t srv_grid_blocks_per_cu(int *blocks)
{
	return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_srv_grid, 64 * SRV_GWAVES, 0);
}

/* Kerr blocks a CU holds beside one FP64 or grid block: every block has a
 * wave on each SIMD, whose 512 registers per lane the waves share
 * (allocated in granules of 8). */
hipError_t srv_kerr_beside_heavy(int *blocks)
{
	hipFuncAttributes k, g, f;
	hipError_t e = hipFuncGetAttributes(&k, reinterpret_cast<const void *>(&k_srv_kerr));
	if (e == hipSuccess)
		e = hipFuncGetAttributes(&g, reinterpret_cast<const void *>(&k_srv_grid));
	if (e == hipSuccess)
		e = hipFuncGetAttributes(&f, reinterpret_cast<const void *>(&k_srv_fb64));
	if (e != hipSuccess)
		return e;
	auto up8 = [](int r) { return (r + 7) / 8 * 8; };
	const int rk = up8(k.numRegs), rh = up8(g.numRegs > f.numRegs ? g.numRegs : f.numRegs);
	*blocks = (rk > 0 && rh > 0 && rh < 512) ? (512 - rh) / rk : 0;
	return hipSuccess;
}

} // namespace gray
