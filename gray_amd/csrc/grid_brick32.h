/*
 * grid_brick32.h -- LDS bricks of the FP32 gridded form: a wave keeps 5x5x5
 * grid points (4x4x4 cells) of the t1 level's Christoffels in LDS and reads
 * the corners of the lanes whose cell lies in it from there; the RHS and
 * RK4 step over it (grid_rhs_brick, grid_rk4_brick).  LDS holds the same
 * FP32 values as global memory and the corner sums keep the global path's
 * order, so a ray's result does not depend on where its corners were read.
 * Used by the render tile kernel and the serving pool's grid kernel.
 */
#pragma once

#include "grid_f32.h"

namespace gray {

constexpr int GB = 5;                       /* brick edge in grid points   */
constexpr int GP_PTS = GB * GB * GB;        /* brick points                */

/* Cell c's 8 corners lie in the brick with origin point o. */
__device__ __forceinline__ bool in_brick5(const int c[3], const int o[3])
{
	return (unsigned)(c[0] - o[0]) <= (unsigned)(GB - 2) &&
	       (unsigned)(c[1] - o[1]) <= (unsigned)(GB - 2) &&
	       (unsigned)(c[2] - o[2]) <= (unsigned)(GB - 2);
}

struct BrickHdr {
	int o[3];                                /* brick origin (point index)  */
	int cool;                                /* evaluations without reload  */
};

/* lev_contract with the corners read from the LDS brick: the 8 corners
 * unrolled, their brick offsets and weights formed once (upper indices
 * clamped as in the stencil: a zero stride means the upper corner is the
 * lower one).  Weights and sums are the stencil's, in its order. */
template <bool PF>
__device__ __forceinline__ void lds_contract(const float4 *__restrict__ brick, const Stencil<float> &st,
                                             const int o[3], const float P[10], float out[4])
{
	const float4 *p0 = brick + ((st.c[0] - o[0]) + GB * (st.c[1] - o[1]) + GB * GB * (st.c[2] - o[2])) * 10;
	const int ox = st.sx ? 10 : 0, oy = st.sy ? 10 * GB : 0, oz = st.sz ? 10 * GB * GB : 0;
	const float gx[2] = {1.f - st.f[0], st.f[0]};
	const float gy[2] = {1.f - st.f[1], st.f[1]};
	const float gz[2] = {1.f - st.f[2], st.f[2]};
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
		out[mu] = 0.f;
	/* PF (the kernels without transfer channels, which have the registers):
	 * corner k + 1's ten 16-byte reads are issued before corner k is
	 * summed, so a wave keeps a corner of LDS latency in flight; the sums
	 * are the same either way */
	float4 cur[10], nxt[10];
	if constexpr (PF) {
#pragma unroll
		for (int i = 0; i < 10; ++i)
			cur[i] = p0[i];
	}
#pragma unroll
	for (int k = 0; k < 8; ++k) {
		const float4 *p = p0 + ((k & 1) ? ox : 0) + ((k & 2) ? oy : 0) + ((k & 4) ? oz : 0);
		if constexpr (PF) {
			if (k < 7) {
				const int k1 = k + 1;
				const float4 *pn = p0 + ((k1 & 1) ? ox : 0) + ((k1 & 2) ? oy : 0) +
				                   ((k1 & 4) ? oz : 0);
#pragma unroll
				for (int i = 0; i < 10; ++i)
					nxt[i] = pn[i];
			}
		}
		const float w = (gx[k & 1] * gy[(k >> 1) & 1]) * gz[k >> 2];
#pragma unroll
		for (int hf = 0; hf < 2; ++hf) {
			float v[20];
#pragma unroll
			for (int i = 0; i < 5; ++i) {
				const float4 q = PF ? cur[5 * hf + i] : p[5 * hf + i];
				v[4 * i] = q.x;
				v[4 * i + 1] = q.y;
				v[4 * i + 2] = q.z;
				v[4 * i + 3] = q.w;
			}
#pragma unroll
			for (int m2 = 0; m2 < 2; ++m2) {
				float tt = 0.f;
#pragma unroll
				for (int c = 0; c < 10; ++c)
					tt += v[10 * m2 + c] * P[c];
				out[2 * hf + m2] += w * tt;
			}
		}
		if constexpr (PF) {
#pragma unroll
			for (int i = 0; i < 10; ++i)
				cur[i] = nxt[i];
		}
	}
}

/* One point's 40 Christoffels into the brick (as load40 reads them). */
template <class S>
__device__ __forceinline__ void brick_store(float4 *brick, int p, const S *src)
{
	float v[40];
	load40(gmem(src), v);
#pragma unroll
	for (int i = 0; i < 10; ++i)
		brick[p * 10 + i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
}

/* Gridded RHS with bricks: grid_rhs_all_f32 with the t1 level's corners
 * from the wave's brick where possible. */
template <int NF, class S>
__device__ __forceinline__ void grid_rhs_brick(const GridView<S> &G, const RtThermalF &rt,
                                              float4 *brick, BrickHdr *hdr, bool use_lds,
                                              const float *s, float *ds, int nf = NF)
{
	Stencil<float> st;
	make_stencil<float, S>(G, s[0], s[1], s[2], s[3], st);
	const float u0 = s[4], u1 = s[5], u2 = s[6], u3 = s[7];
	const float P[10] = {u0 * u0, 2.f * u0 * u1, 2.f * u0 * u2, 2.f * u0 * u3, u1 * u1,
	                     2.f * u1 * u2, 2.f * u1 * u3, u2 * u2, 2.f * u2 * u3, u3 * u3};
	float uGu[4];
	bool inside = false;
	int o[3] = {0, 0, 0};
	if (use_lds) {
		/* the header is wave state: re-read it after the previous
		 * evaluation's writes and take the first active lane's copy, so
		 * every lane makes the same decisions */
		wave_sync();
		o[0] = __builtin_amdgcn_readfirstlane(hdr->o[0]);
		o[1] = __builtin_amdgcn_readfirstlane(hdr->o[1]);
		o[2] = __builtin_amdgcn_readfirstlane(hdr->o[2]);
		const int cool = __builtin_amdgcn_readfirstlane(hdr->cool);
		inside = in_brick5(st.c, o);
		const unsigned long long act = __ballot(1);
		const unsigned long long out = __ballot(!inside);
		const unsigned lane = __lane_id();
		const int leader = __ffsll((long long)act) - 1;
		if (cool > 0) {
			if ((int)lane == leader)
				hdr->cool = cool - 1;
			wave_sync();
		} else if (4 * __popcll(out) > __popcll(act)) {
			/* reload around the first outside ray's cell */
			const int first = __ffsll((long long)out) - 1;
			for (int a = 0; a < 3; ++a) {
				const int c = __builtin_amdgcn_readfirstlane(__shfl(st.c[a], first));
				const int hi_o = G.n[a] - GB;
				o[a] = c - 1 < 0 ? 0 : (c - 1 > hi_o ? hi_o : c - 1);
			}
			wave_sync();
			const int n_act = __popcll(act);
			const int r = __popcll(act & ((1ull << lane) - 1ull));
			const uint32_t nx = (uint32_t)G.n[0], nxy = nx * (uint32_t)G.n[1];
			for (int p = r; p < GP_PTS; p += n_act) {
				const uint32_t pt = (uint32_t)(o[2] + p / (GB * GB)) * nxy +
				                    (uint32_t)(o[1] + (p / GB) % GB) * nx +
				                    (uint32_t)(o[0] + p % GB);
				brick_store<S>(brick, p, G.gam[0] + (size_t)pt * 40);
			}
			inside = in_brick5(st.c, o);
			const int still = __popcll(__ballot(!inside));
			if ((int)lane == leader) {
				hdr->o[0] = o[0];
				hdr->o[1] = o[1];
				hdr->o[2] = o[2];
				hdr->cool = 2 * still > n_act ? 8 : 0;
			}
			wave_sync();
		}
	}
	if (inside)
		lds_contract<NF == 0>(brick, st, o, P, uGu);
	else if constexpr (NF == 0)
		lev_contract_pf<S>(G.gam[0], st, P, uGu);
	else
		lev_contract<S>(G.gam[0], st, P, uGu);
	if (st.wt > 0.f) {
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
		if (nf > 0) {
			float kt, ne, bkcos;
			flow_f32<S>(G, rt, st, u, kt, ne, bkcos);
			thermal_f32<NF>(rt, ne, -kt, bkcos, s + 8 + NF, ds + 8, ds + 8 + NF, nf);
		} else {
#pragma unroll
			for (int i = 8; i < 8 + 2 * NF; ++i)
				ds[i] = 0.f;
		}
	}
}

/* grid_rk4_f32c over grid_rhs_brick. */
template <int NF, class S>
__device__ __forceinline__ void grid_rk4_brick(const GridView<S> &G, const RtThermalF &rt,
                                              float4 *brick, BrickHdr *hdr, bool use_lds,
                                              float *hi, float *lo, float dt, int nf = NF)
{
	constexpr int ND = 8 + 2 * NF;
	float k[ND], acc[ND], x[ND];
	const float hdt = 0.5f * dt;
	grid_rhs_brick<NF, S>(G, rt, brick, hdr, use_lds, hi, k, nf);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] = k[i];
		x[i] = hi[i] + hdt * k[i];
	}
	grid_rhs_brick<NF, S>(G, rt, brick, hdr, use_lds, x, k, nf);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + hdt * k[i];
	}
	grid_rhs_brick<NF, S>(G, rt, brick, hdr, use_lds, x, k, nf);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		acc[i] += 2.0f * k[i];
		x[i] = hi[i] + dt * k[i];
	}
	grid_rhs_brick<NF, S>(G, rt, brick, hdr, use_lds, x, k, nf);
	const float dt6 = dt * (1.0f / 6.0f);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		float y = (acc[i] + k[i]) * dt6 + lo[i];
		float tt = hi[i] + y;
		lo[i] = y - (tt - hi[i]);
		hi[i] = tt;
	}
}

} // namespace gray
