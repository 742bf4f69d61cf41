/*
 * kernels_grid64.hip -- plain FP64 form of the numerical-spacetime path:
 * camera initial conditions with the interpolated metric, the tile kernel
 * (every dump interval of a render in one launch, LDS-staged stencils),
 * and g(u,u) diagnostics.
 *
 * Built with -ffp-contract=off and evaluated in the order of the FP64 oracle
 * (interpolate each Christoffel, blend in time, then contract as
 * sim-org/dyst.cl:186-191 does), so on uniform grids it reproduces the
 * oracle bit for bit.
 */
#include <climits>
#include <cstdlib>

#include "grid_f64.h"
#include "launch.h"

namespace gray {

/* ------------------------------------------------------------------
 * Tile kernel of the plain FP64 gridded form with LDS bricks.
 *
 * Each wave holds one tile of 8x8 rays (dealt to the XCDs in patches, see
 * tile_of in grid_device.h) and runs all dump
 * intervals in one flat loop: each pass is one RK4 step of every ray not
 * yet stopped, and a stopped ray leaves the loop for good.  So every lane
 * still in the loop takes part in every RHS, and the brick's origin can be
 * kept in registers: the active lanes compute it from the same ballots.
 * The brick holds the t1 level's Christoffels of 4x4x4 points in their
 * stored type (64 x 40 x sizeof(S) bytes per wave); a corner read from it
 * is the same value as from global memory, and the corner sums keep the
 * oracle's order, so the results stay bit for bit the oracle's.
 * ------------------------------------------------------------------ */

constexpr int GT_WAVES = 2;                 /* waves per block */

/* sample_lev_f64 with the corners in the brick (same sums, same order).
 * Unrolled, with corner k + 1's reads issued before corner k is summed, so
 * a wave keeps a corner of LDS latency in flight (253 VGPRs, two waves per
 * SIMD); the global-memory corner loop stays rolled (unrolled, its
 * registers cost more than its overlap gains). */
template <int NC, class S>
__device__ __forceinline__ void sample_brick_f64(const S *__restrict__ brick, int base,
                                                 const Stencil<double> &st, const int o[3],
                                                 double out[NC])
{
	const int p0 = (st.c[0] - o[0]) + 4 * (st.c[1] - o[1]) + 16 * (st.c[2] - o[2]);
	const int ox = st.sx ? 1 : 0, oy = st.sy ? 4 : 0, oz = st.sz ? 16 : 0;
	S cur[NC], nxt[NC];
#pragma unroll
	for (int c = 0; c < NC; ++c)
		cur[c] = brick[p0 * 40 + base + c];
#pragma unroll
	for (int k = 0; k < 8; ++k) {
		if (k < 7) {
			const int k1 = k + 1;
			const int pn = p0 + ((k1 & 1) ? ox : 0) + ((k1 & 2) ? oy : 0) + ((k1 & 4) ? oz : 0);
#pragma unroll
			for (int c = 0; c < NC; ++c)
				nxt[c] = brick[pn * 40 + base + c];
		}
		const double w = st.weight(k);
#pragma unroll
		for (int c = 0; c < NC; ++c)
			out[c] = k == 0 ? w * (double)cur[c] : out[c] + w * (double)cur[c];
#pragma unroll
		for (int c = 0; c < NC; ++c)
			cur[c] = nxt[c];
	}
}

/* Keeps the wave's brick around its rays: when more than a quarter of the
 * active rays' cells lie outside it, reload it around the first such cell
 * (one point per active lane); if half are still outside, wait 8
 * evaluations before the next try.  Returns whether this lane's cell is in
 * the brick. */
template <class S>
__device__ __forceinline__ bool brick_update_f64(const GridView<S> &G, S *brick, BrickState &b,
                                                 const Stencil<double> &st)
{
	bool inside = in_brick(st.c, b.o);
	const unsigned long long act = __ballot(1);
	const unsigned long long out = __ballot(!inside);
	if (b.cool > 0) {
		--b.cool;
		return inside;
	}
	if (!(4 * __popcll(out) > __popcll(act)))
		return inside;
	const int first = __ffsll((long long)out) - 1;
	for (int a = 0; a < 3; ++a) {
		const int c = __shfl(st.c[a], first);
		const int hi = G.n[a] - 4;
		b.o[a] = c - 1 < 0 ? 0 : (c - 1 > hi ? hi : c - 1);
	}
	wave_sync();
	const unsigned lane = __lane_id();
	const int n_act = __popcll(act);
	const int r = __popcll(act & ((1ull << lane) - 1ull));
	const uint32_t nx = (uint32_t)G.n[0], nxy = nx * (uint32_t)G.n[1];
	for (int p = r; p < 64; p += n_act) {
		const uint32_t pt = (uint32_t)(b.o[2] + (p >> 4)) * nxy +
		                    (uint32_t)(b.o[1] + ((p >> 2) & 3)) * nx + (uint32_t)(b.o[0] + (p & 3));
		const S *src = G.gam[0] + (size_t)pt * 40;
		S *dst = brick + p * 40;
#pragma unroll
		for (int c = 0; c < 40; ++c)
			dst[c] = src[c];
	}
	wave_sync();
	inside = in_brick(st.c, b.o);
	const int still = __popcll(__ballot(!inside));
	b.cool = 2 * still > n_act ? 8 : 0;
	return inside;
}

template <int NF, class S>
__device__ __forceinline__ void grid_rhs_brick_f64(const GridView<S> &G, const RtThermal &rt,
                                                   S *brick, BrickState &b, bool use_lds,
                                                   const double *s, double *ds)
{
	Stencil<double> st;
	make_stencil<double, S>(G, s[0], s[1], s[2], s[3], st);
	const bool inside = use_lds && brick_update_f64<S>(G, brick, b, st);
	ds[0] = s[4]; ds[1] = s[5]; ds[2] = s[6]; ds[3] = s[7];
	const double u[4] = {s[4], s[5], s[6], s[7]};
	double uGu[4];
#pragma unroll 1
	for (int mu = 0; mu < 4; ++mu) {
		double g[10], v[4];
		if (inside)
			sample_brick_f64<10, S>(brick, 10 * mu, st, b.o, g);
		else
			sample_lev_f64<10, 40, S>(G.gam[0], 10 * mu, st, g);
		if (st.wt > 0.0) {
			double y2[10];
			sample_lev_f64<10, 40, S>(G.gam[1], 10 * mu, st, y2);
#pragma unroll
			for (int c = 0; c < 10; ++c)
				g[c] = g[c] + st.wt * (y2[c] - g[c]);
		}
#pragma unroll
		for (int a = 0; a < 4; ++a)
			v[a] = ((g[sym(a, 0)] * u[0] + g[sym(a, 1)] * u[1]) + g[sym(a, 2)] * u[2]) +
			       g[sym(a, 3)] * u[3];
		const double r = ((u[0] * v[0] + u[1] * v[1]) + u[2] * v[2]) + u[3] * v[3];
		uGu[0] = mu == 0 ? r : uGu[0];
		uGu[1] = mu == 1 ? r : uGu[1];
		uGu[2] = mu == 2 ? r : uGu[2];
		uGu[3] = mu == 3 ? r : uGu[3];
	}
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
		ds[4 + mu] = -uGu[mu] + uGu[0] * u[mu];
	if constexpr (NF > 0) {
		double m[4], ne;
		sample_f64<4, 10, S>(G.met, 0, st, m);
		sample_f64<1, 1, S>(G.rho, 0, st, &ne);
		const double kt = ((m[0] * s[4] + m[1] * s[5]) + m[2] * s[6]) + m[3] * s[7];
		double bkcos = 0.0;
		if (rt.b > 0.0) {
			/* k_z from the z row {3, 6, 8, 9} of the metric; components 6..9 */
			double mz[4];
			sample_f64<4, 10, S>(G.met, 6, st, mz);
			const double kz = ((m[3] * s[4] + mz[0] * s[5]) + mz[2] * s[6]) + mz[3] * s[7];
			bkcos = (kz * rt.b) / ((-kt) * rt.b + RT_EPS);
		}
		thermal_f64<NF>(rt, ne, -kt, bkcos, s + 8 + NF, ds + 8, ds + 8 + NF);
	}
}

/* sim-org/RK4.cl:37-47 on the whole state (8 + 2 NF components). */
template <int NF, class S>
__device__ __forceinline__ void grid_rk4_brick_f64(const GridView<S> &G, const RtThermal &rt,
                                                   S *brick, BrickState &b, bool use_lds, double *s,
                                                   double dt)
{
	constexpr int ND = 8 + 2 * NF;
	double k1[ND], k2[ND], k[ND], x[ND];
	const double hdt = 0.5 * dt;
	grid_rhs_brick_f64<NF, S>(G, rt, brick, b, use_lds, s, k1);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = s[i] + hdt * k1[i];
	grid_rhs_brick_f64<NF, S>(G, rt, brick, b, use_lds, x, k2);
#pragma unroll
	for (int i = 0; i < ND; ++i) x[i] = s[i] + hdt * k2[i];
	grid_rhs_brick_f64<NF, S>(G, rt, brick, b, use_lds, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		k2[i] = k2[i] + k[i];
		x[i] = s[i] + dt * k[i];
	}
	grid_rhs_brick_f64<NF, S>(G, rt, brick, b, use_lds, x, k);
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = s[i] + div6(dt * ((k1[i] + 2.0 * k2[i]) + k[i]));
}

template <int NF, class S>
__global__ __launch_bounds__(64 * GT_WAVES) void k_grid_tiles_f64(GridView<S> G, RtThermal rt,
                                                                  StepRule sr,
                                                                  double *__restrict__ data,
                                                                  double *__restrict__ info,
                                                                  int64_t *__restrict__ steps,
                                                                  size_t n_rays, int W, int H,
                                                                  int bricks)
{
	constexpr int ND = 8 + 2 * NF;
	__shared__ S s_brick[GT_WAVES][64 * 40];
	const unsigned lane = __lane_id();
	const int wv = (int)(threadIdx.x >> 6);
	const unsigned tw = (unsigned)(W + 7) / 8u, th = (unsigned)(H + 7) / 8u;
	unsigned tx, ty;
	if (!tile_of(blockIdx.x, (unsigned)wv, GT_WAVES, tw, th, (bricks & 2) != 0, tx, ty))
		return;
	const unsigned x = tx * 8u + (lane & 7u), y = ty * 8u + (lane >> 3);
	if (x >= (unsigned)W || y >= (unsigned)H)
		return;
	const size_t g = (size_t)y * (size_t)W + x;
	S *brick = s_brick[wv];
	BrickState b = {{-1000000, -1000000, -1000000}, 0};   /* no cell matches: first RHS loads */
	const bool use_lds = (bricks & 1) && G.n[0] >= 4 && G.n[1] >= 4 && G.n[2] >= 4;

	double s[ND];
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = data[(size_t)i * n_rays + g];
	int64_t nst = 0;
	int h = 0, d = 0;
	for (;;) {
		double xs = grid_getdt<double, S>(G, s[1], s[2], s[3], s[5], s[6], s[7], sr.dts) / sr.dt *
		            (double)sr.n;
		int dh = (xs == xs) ? (int)xs : 0;
		if (!dh)
			break;
		if (dh > sr.n - h)
			dh = sr.n - h;
		grid_rk4_brick_f64<NF, S>(G, rt, brick, b, use_lds, s, (double)dh * sr.dt / (double)sr.n);
		++nst;
		h += dh;
		if (h >= sr.n) {
			h = 0;
			if (++d >= sr.n_dump)
				break;
		}
	}
#pragma unroll
	for (int i = 0; i < ND; ++i)
		data[(size_t)i * n_rays + g] = s[i];
	info[g] = grid_getuu_f64<S>(G, s);
	steps[g] += nst;
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

/* Covariant view: out[g][8] = (q, k_mu), k = g(q) u with the interpolated
 * metric. */
template <class S>
__global__ __launch_bounds__(256) void k_grid_covariant(GridView<S> G, const double *__restrict__ data,
                                                        double *__restrict__ out, size_t n_rays)
{
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	double s[8];
#pragma unroll
	for (int i = 0; i < 8; ++i)
		s[i] = data[(size_t)i * n_rays + g];
	double k[4];
	grid_down_f64<S>(G, s, k);
	double *o = out + (size_t)8 * g;
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		o[i] = s[i];
		o[4 + i] = k[i];
	}
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
static void grid_tiles_f64_nf(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                              double *data, double *info, int64_t *steps, size_t n_rays, int W,
                              int H, int bricks, unsigned nb, hipStream_t st)
{
	if (gd.dtype == 4)
		hipLaunchKernelGGL((k_grid_tiles_f64<NF, float>), dim3(nb), dim3(64 * GT_WAVES), 0, st,
		                   gd.view<float>(), rt, sr, data, info, steps, n_rays, W, H, bricks);
	else
		hipLaunchKernelGGL((k_grid_tiles_f64<NF, double>), dim3(nb), dim3(64 * GT_WAVES), 0, st,
		                   gd.view<double>(), rt, sr, data, info, steps, n_rays, W, H, bricks);
}

hipError_t launch_grid_tiles_f64(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                                 double *data, double *info, int64_t *steps, size_t n_rays, int W,
                                 int H, bool bricks, hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
	if (W <= 0 || H <= 0 || (size_t)W * (size_t)H != n_rays)
		return hipErrorInvalidValue;
	const size_t tiles = (((size_t)W + 7) / 8) * (((size_t)H + 7) / 8);
	if (tiles > (size_t)INT_MAX / 2)
		return hipErrorInvalidValue;
	/* tiles in XCD patches (grid_device.h: tile_of); measured on 512^2 rays
	 * through the 128^3 grid: FP32 form 300 -> 287 ms, FP64 800 -> 737 ms */
	const bool xcd = true;
	const size_t nb = tile_blocks((unsigned)(W + 7) / 8u, (unsigned)(H + 7) / 8u, GT_WAVES, xcd);
#define L(NF) grid_tiles_f64_nf<NF>(gd, rt, sr, data, info, steps, n_rays, W, H, \
	                            (bricks ? 1 : 0) | (xcd ? 2 : 0), \
	                            (unsigned)nb, st)
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

hipError_t launch_grid_covariant(const GridDesc &gd, const double *data, double *out, size_t n_rays,
                                 hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
	if (gd.dtype == 4)
		hipLaunchKernelGGL(k_grid_covariant<float>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st,
		                   gd.view<float>(), data, out, n_rays);
	else
		hipLaunchKernelGGL(k_grid_covariant<double>, dim3(blocks_for(n_rays, 256)), dim3(256), 0,
		                   st, gd.view<double>(), data, out, n_rays);
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
