/*
 * grid_device.h -- per-ray device math of the numerical-spacetime path:
 * cell location and trilinear-in-space / linear-in-time interpolation of the
 * gridded metric and Christoffel symbols (sim-org/interp.cl:53-205) and the
 * coordinate-time geodesic RHS of sim-org/dyst.cl:112-194.
 *
 * HBM layout (per time level): Christoffels interleaved per grid point,
 * gam[p*40 + 10*mu + c] with c in tt tx ty tz xx xy xz yy yz zz, and the
 * metric met[p*10 + c], and the density rho[p]; p = (k*ny + j)*nx + i.
 * One RHS reads the 8 corner
 * points' 40 values (one contiguous 160 B / 320 B run each) from both
 * levels; the cell and its weights are computed once per RHS for all
 * fields (the reference recomputes the map for each of its 102 image reads).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fpdiv.h"
#include "grayhip_params.h"

namespace gray {

/* Global-memory pointers.  The grid's level pointers reach the serving
 * kernels through a request table in memory, where the compiler cannot
 * tell their address space and would read through flat loads (which also
 * count against LDS waits); the grid readers load through these instead. */
template <class T>
using gptr = const __attribute__((address_space(1))) T *;

template <class T>
__device__ __forceinline__ gptr<T> gmem(const T *p)
{
	return (gptr<T>)p;
}

template <class S>
struct GridView {
	const S *gam[2];      /* per level: [p][40]                        */
	const S *met[2];      /* per level: [p][10]                        */
	const S *rho[2];      /* per level: [p], density (may be null)     */
	int    n[3];
	int    map;           /* GRAYHIP_MAP_*                             */
	double lo[3], hi[3];
	double dx[3];         /* (hi - lo) / (n - 1), uniform map           */
	double idx[3];        /* (n - 1) / (hi - lo), uniform map           */
	double fa[3], fb[3];  /* fisheye A, B per axis                      */
	double t1, t2;        /* level times; equal: space only             */
	double t_off;         /* grid time at q^t = 0                       */
	int    raw_icond;     /* 1: no null normalisation                   */
	int    ref_rule;      /* 1: no box-exit stop                        */
};

/* Symmetric index of (a, b) among the 10 unique components. */
__device__ __forceinline__ int sym(int a, int b)
{
	/* rows: {0,1,2,3},{1,4,5,6},{2,5,7,8},{3,6,8,9} */
	const int lo_ = a < b ? a : b, hi_ = a < b ? b : a;
	return lo_ == 0 ? hi_ : (lo_ == 1 ? 3 + hi_ : (lo_ == 2 ? 5 + hi_ : 9));
}

/* The 8 corners of a cell: corner k (bits x, y, z of k) is point
 * base + (k&1 ? sx : 0) + (k&2 ? sy : 0) + (k&4 ? sz : 0) with weight
 * (gx * gy) * gz, g = f or 1 - f per axis; the upper index is clamped to
 * the last point (sx, sy, sz = 0 there). */
template <class T> struct Stencil {
	uint32_t base, sx, sy, sz;
	int c[3];        /* lower corner (cell) indices i0, j0, k0 */
	T f[3];
	T wt;            /* time weight; < 0: space only */

	__device__ __forceinline__ uint32_t corner(int k) const
	{
		return base + ((k & 1) ? sx : 0u) + ((k & 2) ? sy : 0u) + ((k & 4) ? sz : 0u);
	}
	__device__ __forceinline__ T weight(int k) const
	{
		const T gx = (k & 1) ? f[0] : T(1) - f[0];
		const T gy = (k & 2) ? f[1] : T(1) - f[1];
		const T gz = (k & 4) ? f[2] : T(1) - f[2];
		return (gx * gy) * gz;
	}
};

template <class T>
__device__ __forceinline__ T ex_floor(T x) { return floor(x); }

/* interp.cl:69-164 for one axis in the arithmetic of T. */
template <class T, class S>
__device__ __forceinline__ void locate(const GridView<S> &G, int ax, T x, int &i0, T &f)
{
	const int n = G.n[ax];
	T xi, x0, x1;
	int i;
	if (G.map == GRAYHIP_MAP_FISHEYE) {
		const T A = T(G.fa[ax]), B = T(G.fb[ax]);
		xi = ex_floor((cbrt(asinh(x)) - B) / A);
		i = xi >= T(0) ? (xi <= T(n - 1) ? (int)xi : n - 1) : 0;
		T c0 = A * T(i) + B, c1 = A * T(i + 1) + B;
		x0 = sinh((c0 * c0) * c0);
		x1 = sinh((c1 * c1) * c1);
	} else if constexpr (sizeof(T) == 4) {
		/* FP32 form: on a uniform axis i + (x - x0) / (x1 - x0) is
		 * (x - lo) / dx, formed with the host's 1/dx (one multiply; the
		 * plain FP64 form keeps the reference's two divisions) */
		T u = (x - T(G.lo[ax])) * T(G.idx[ax]);
		if (!(u > T(0)))
			u = T(0);
		if (u > T(n - 1))
			u = T(n - 1);
		T fl = ex_floor(u);
		i0 = (int)fl;
		f = u - fl;
		return;
	} else {
		const T lo = T(G.lo[ax]), dx = T(G.dx[ax]);
		xi = ex_floor((x - lo) / dx);
		i = xi >= T(0) ? (xi <= T(n - 1) ? (int)xi : n - 1) : 0;
		x0 = lo + T(i) * dx;
		x1 = lo + T(i + 1) * dx;
	}
	T u = T(i) + (x - x0) / (x1 - x0);
	if (!(u > T(0)))
		u = T(0);
	if (u > T(n - 1))
		u = T(n - 1);
	T fl = ex_floor(u);
	i0 = (int)fl;
	f = u - fl;
}

template <class T, class S>
__device__ __forceinline__ void make_stencil(const GridView<S> &G, T qt, T qx, T qy, T qz,
                                             Stencil<T> &st)
{
	int i[3];
	T f[3];
	locate<T, S>(G, 0, qx, i[0], f[0]);
	locate<T, S>(G, 1, qy, i[1], f[1]);
	locate<T, S>(G, 2, qz, i[2], f[2]);
	const uint32_t nx = (uint32_t)G.n[0], nxy = nx * (uint32_t)G.n[1];
	st.base = (uint32_t)i[2] * nxy + (uint32_t)i[1] * nx + (uint32_t)i[0];
	st.sx = i[0] + 1 < G.n[0] ? 1u : 0u;
	st.sy = i[1] + 1 < G.n[1] ? nx : 0u;
	st.sz = i[2] + 1 < G.n[2] ? nxy : 0u;
	st.c[0] = i[0];
	st.c[1] = i[1];
	st.c[2] = i[2];
	st.f[0] = f[0];
	st.f[1] = f[1];
	st.f[2] = f[2];
	if (G.t1 == G.t2) {
		st.wt = T(-1);
	} else {
		T w = ((T(G.t_off) + qt) - T(G.t1)) / (T(G.t2) - T(G.t1));
		st.wt = w < T(0) ? T(0) : (w > T(1) ? T(1) : w);
	}
}

/* Tile (8x8 rays) of wave wv of block b, in a grid of tw x th tiles.
 * Row-major order (xcd_patches = false): tile waves b + wv.  XCD order:
 * blocks reach the 8 XCDs round robin (block b on XCD b mod 8), each XCD
 * with its own L2; tiles are grouped in patches of TILE_PATCH x TILE_PATCH
 * and patch c is dealt to XCD c mod 8, so that rays which share grid cells
 * share one L2, while the patches an XCD holds lie all over the image (an
 * even share of the costly rays near the hole). */
constexpr unsigned TILE_PATCH = 4;

__host__ __device__ inline unsigned tile_blocks(unsigned tw, unsigned th, unsigned waves,
                                                bool xcd_patches)
{
	if (!xcd_patches)
		return (tw * th + waves - 1) / waves;
	const unsigned np = ((tw + TILE_PATCH - 1) / TILE_PATCH) * ((th + TILE_PATCH - 1) / TILE_PATCH);
	return ((np + 7) / 8) * 8 * (TILE_PATCH * TILE_PATCH / waves);
}

__host__ __device__ inline bool tile_of(unsigned b, unsigned wv, unsigned waves, unsigned tw,
                                        unsigned th, bool xcd_patches, unsigned &tx, unsigned &ty)
{
	if (!xcd_patches) {
		const unsigned t = b * waves + wv;
		tx = t % tw;
		ty = t / tw;
		return t < tw * th;
	}
	const unsigned ptw = (tw + TILE_PATCH - 1) / TILE_PATCH;
	const unsigned np = ptw * ((th + TILE_PATCH - 1) / TILE_PATCH);
	const unsigned bpp = TILE_PATCH * TILE_PATCH / waves;   /* blocks per patch */
	const unsigned k = b / 8u;
	const unsigned c = (k / bpp) * 8u + b % 8u;
	if (c >= np)
		return false;
	const unsigned t = (k % bpp) * waves + wv;
	tx = (c % ptw) * TILE_PATCH + t % TILE_PATCH;
	ty = (c / ptw) * TILE_PATCH + t / TILE_PATCH;
	return tx < tw && ty < th;
}

/* LDS bricks: a wave's copy of 4x4x4 grid points (3x3x3 cells) from
 * origin o; a cell c lies in it when 0 <= c - o <= 2 on every axis. */
__device__ __forceinline__ bool in_brick(const int c[3], const int o[3])
{
	return (unsigned)(c[0] - o[0]) <= 2u && (unsigned)(c[1] - o[1]) <= 2u &&
	       (unsigned)(c[2] - o[2]) <= 2u;
}

/* A wave's brick state: origin point and evaluations left before the next
 * reload may be tried.  Kept in registers by kernels whose lanes all take
 * part in every RHS, so that every active lane holds the same values. */
struct BrickState {
	int o[3];
	int cool;
};

/* Orders a wave's LDS writes before its later reads (and reads before
 * later writes): wave-scope fences around a scheduling barrier. */
__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Coordinate speed |dx/dt|^2 above which a ray counts as captured: the
 * horizon proxy of the gridded path.  Outside a horizon light moves at
 * |dx/dt| of order 1; deep inside, near the singularity, the interpolated
 * Christoffels make the speed grow until RK4 overflows. */
#define GRID_U_CAP2 16.0

/* Box-exit stop, the horizon proxy and the fixed 0.01 cap of dyst
 * (getrr = geteps = 1); u^t = 1 in this coordinate-time form. */
template <class T, class S>
__device__ __forceinline__ T grid_getdt(const GridView<S> &G, T qx, T qy, T qz, T ux, T uy,
                                        T uz, T dt)
{
	if (!G.ref_rule) {
		if (!(qx >= T(G.lo[0]) && qx <= T(G.hi[0])) || !(qy >= T(G.lo[1]) && qy <= T(G.hi[1])) ||
		    !(qz >= T(G.lo[2]) && qz <= T(G.hi[2])))
			return T(0);
		const T v2 = (ux * ux + uy * uy) + uz * uz;
		if (!(v2 <= T(GRID_U_CAP2)))
			return T(0);
	}
	const T cap = (T(0.01) * T(1)) * T(1);
	if (fabs(dt) > cap)
		return dt > T(0) ? cap : -cap;
	return dt;
}

} // namespace gray

namespace gray {

/* Host-side description of a resident spacetime (two level slots). */
struct GridDesc {
	int dtype;                    /* 4 or 8 bytes per stored real          */
	const void *gam[2], *met[2];  /* device arrays of the t1 / t2 slots     */
	const void *rho[2];           /* densities of the slots (may be null)  */
	int n[3];
	int map;
	double lo[3], hi[3], dx[3], idx[3], fa[3], fb[3];
	double t1, t2, t_off;
	int raw_icond, ref_rule;

	template <class S> GridView<S> view() const
	{
		GridView<S> v;
		for (int l = 0; l < 2; ++l) {
			v.gam[l] = (const S *)gam[l];
			v.met[l] = (const S *)met[l];
			v.rho[l] = (const S *)rho[l];
		}
		for (int a = 0; a < 3; ++a) {
			v.n[a] = n[a];
			v.lo[a] = lo[a];
			v.hi[a] = hi[a];
			v.dx[a] = dx[a];
			v.idx[a] = idx[a];
			v.fa[a] = fa[a];
			v.fb[a] = fb[a];
		}
		v.map = map;
		v.t1 = t1;
		v.t2 = t2;
		v.t_off = t_off;
		v.raw_icond = raw_icond;
		v.ref_rule = ref_rule;
		return v;
	}
};

} // namespace gray
