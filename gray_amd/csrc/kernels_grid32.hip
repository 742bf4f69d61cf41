/*
 * kernels_grid32.hip -- FP32 form of the numerical-spacetime path: a
 * tile kernel, FP32 RHS and RK4 stages on a compensated (hi + lo)
 * state, FP32 or FP64 grid storage; transfer channels (the reference flow)
 * on the state scaled by K as described in rt_thermal.h.
 *
 * The RHS contracts each corner's 40 Christoffels with the ten products
 * u^a u^b first and interpolates the four results (linear operations
 * commute), so a corner costs one contiguous 160 B (FP32) or 320 B (FP64)
 * read and 40 FMAs, and no per-field interpolation state is kept.
 */
#include <climits>
#include <cstdlib>

#include "grid_brick32.h"
#include "launch.h"

namespace gray {

/* ------------------------------------------------------------------
 * Tile kernel of the FP32 gridded form, with LDS-staged interpolation
 * stencils.
 *
 * A wave holds an 8x8 tile of the image (a bundle of neighbouring rays)
 * and runs every dump interval of its rays in registers; a ray that ends
 * leaves the wave's active mask, and the wave's decisions below are
 * ballots over the rays still active.  Every wave keeps a brick of 5x5x5
 * grid points (4x4x4 cells) of the t1 level's Christoffels in LDS
 * (125 x 160 B = 20 KB; four waves per block, two blocks per CU at the
 * kernel's two waves per SIMD); a lane whose 8 corners lie in the brick
 * reads them from LDS, the others from global memory.  When more than a
 * quarter of the wave's rays have left the brick, the wave reloads it
 * around the cell of its first such ray (about two points per lane); if
 * that still leaves half of them outside (rays spread near the hole),
 * reloading pauses for 8 evaluations.  The brick's origin lives in LDS with it.  LDS holds the
 * same FP32 values as global memory and the contraction order is the
 * global path's, so a ray's result does not depend on where its corners
 * were read (bricks switched off give the same bits).
 * ------------------------------------------------------------------ */

constexpr int GP_WAVES = 4;                 /* waves per block             */
/* One launch = n_dump intervals of every ray; each wave holds one tile of
 * 8x8 rays, tiles dealt to the XCDs in 4x4-tile patches (tile_of,
 * grid_device.h), so neighbouring rays share an L2 while every XCD gets
 * patches from all over the image.  All resident waves start together
 * and advance at similar rates, so the rays in flight sweep the grid as one
 * front and the cells they read stay in L2. */
template <int NF, class S>
/* Two waves per SIMD (as the LDS bricks allow) up to 5 channels, even where
 * that spills (3 to 5 channels: 52-196 B/lane, 1.6x faster than one wave on
 * 512^2 rays through the 128^3 grid, same bits); from 6 channels the spills
 * grow past 1 KB/lane (3.3x slower at 8), and one wave holds the state. */
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NF <= 5 ? 2 : 1)))
void k_grid_tiles_f32(GridView<S> G, RtThermal rt64, StepRule sr, double *__restrict__ data,
                      int64_t *__restrict__ steps, size_t n_rays, int W, int H, int bricks,
                      unsigned long long *__restrict__ n_steps)
{
	constexpr int ND = 8 + 2 * NF;
	__shared__ float4 s_brick[GP_WAVES][GP_PTS * 10];
	__shared__ BrickHdr s_hdr[GP_WAVES];
	const unsigned lane = __lane_id();
	const int wv = (int)(threadIdx.x >> 6);
	float4 *brick = s_brick[wv];
	BrickHdr *hdr = &s_hdr[wv];
	if (lane == 0) {
		/* an origin no cell matches: the first evaluation loads */
		hdr->o[0] = hdr->o[1] = hdr->o[2] = -1000000;
		hdr->cool = 0;
	}
	wave_sync();

	const unsigned tw = (unsigned)(W + 7) / 8u, th = (unsigned)(H + 7) / 8u;
	unsigned tx, ty;
	if (!tile_of(blockIdx.x, (unsigned)wv, GP_WAVES, tw, th, (bricks & 2) != 0, tx, ty))
		return;
	const unsigned x = tx * 8u + (lane & 7u), y = ty * 8u + (lane >> 3);
	if (x >= (unsigned)W || y >= (unsigned)H)
		return;
	const size_t g = (size_t)y * (size_t)W + x;

	const bool use_lds = (bricks & 1) && G.n[0] >= GB && G.n[1] >= GB && G.n[2] >= GB;
	const RtThermalF rt = rt_thermal_f32(rt64);
	const double inv_scale = NF > 0 ? 1.0 / rt64.scale : 1.0;
	float hi[ND], lo[ND];
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		double v = data[(size_t)i * n_rays + g];
		if (i >= 8)
			v *= inv_scale;
		hi[i] = (float)v;
		lo[i] = (float)(v - (double)hi[i]);
	}
	int64_t st = steps[g];
	const int64_t st0 = st;
	const float dts = (float)sr.dts;
	const double dt_over_n = sr.dt / (double)sr.n;
	const double n_over_dt = (double)sr.n / sr.dt;   /* gdt n / dt as one multiply */
	for (int d = 0; d < sr.n_dump; ++d) {
		int dh;
		bool retired = false;
		for (int h = 0; h < sr.n; h += dh) {
			float gdt = grid_getdt<float, S>(G, hi[1], hi[2], hi[3], hi[5], hi[6], hi[7], dts);
			double xd = (double)gdt * n_over_dt;
			dh = (xd == xd) ? (int)xd : 0;
			if (!dh) {
				retired = true;
				break;
			}
			if (dh > sr.n - h)
				dh = sr.n - h;
			grid_rk4_brick<NF, S>(G, rt, brick, hdr, use_lds, hi, lo,
			                     (float)((double)dh * dt_over_n));
			++st;
		}
		if (retired)
			break;
	}
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		double v = (double)hi[i] + (double)lo[i];
		if (i >= 8)
			v *= rt64.scale;
		data[(size_t)i * n_rays + g] = v;
	}
	steps[g] = st;
	atomicAdd(n_steps, (unsigned long long)(st - st0));
}

template <int NF, class S>
static void grid_tiles_launch(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                              double *data, int64_t *steps, size_t n_rays, int W, int H,
                              int bricks, unsigned long long *n_steps, unsigned n_blocks,
                              hipStream_t st)
{
	hipLaunchKernelGGL((k_grid_tiles_f32<NF, S>), dim3(n_blocks), dim3(256), 0, st,
	                   gd.view<S>(), rt, sr, data, steps, n_rays, W, H, bricks, n_steps);
}

hipError_t launch_grid_tiles_f32(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                                 double *data, int64_t *steps, size_t n_rays, int W, int H,
                                 bool bricks, unsigned long long *n_steps, hipStream_t st)
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
	const size_t nb = tile_blocks((unsigned)(W + 7) / 8u, (unsigned)(H + 7) / 8u, GP_WAVES, xcd);
	const int b = (bricks ? 1 : 0) | (xcd ? 2 : 0);
	if (gd.dtype == 4) {
#define L(NF) grid_tiles_launch<NF, float>(gd, rt, sr, data, steps, n_rays, W, H, b, n_steps, \
	                                   (unsigned)nb, st)
		GRAY_NF_SWITCH(rt.n_freq, L)
#undef L
	} else {
#define L(NF) grid_tiles_launch<NF, double>(gd, rt, sr, data, steps, n_rays, W, H, b, n_steps, \
	                                    (unsigned)nb, st)
		GRAY_NF_SWITCH(rt.n_freq, L)
#undef L
	}
	return hipGetLastError();
}

} // namespace gray
