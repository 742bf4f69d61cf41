/*
 * launch.h -- host-side launchers of the gfx950 kernels (one per kernel
 * family), called by the C-ABI layer.
 */
#pragma once

#include <hip/hip_runtime.h>

#include "grayhip_params.h"
#include "grid_device.h"
#include "rt_thermal.h"

namespace gray {

/* Instantiate CALL(n) for the run-time channel count NFV (0..8) inside a
 * launcher returning hipError_t. */
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
                        size_t n_rays, hipStream_t st);
hipError_t launch_evolve_f64(const RayParams &p, const StepRule &sr, double *data, double *info,
                             int64_t *steps, size_t n_rays, hipStream_t st);
hipError_t launch_fallback_f64(const RayParams &p, const StepRule &sr, double *data, double *info,
                               int64_t *steps, size_t n_rays, const Fallback *fb,
                               Counters *ctr, unsigned max_fb, hipStream_t st);
hipError_t launch_soa_to_aos(const double *soa, void *aos, int real_size, size_t n_rays, int nd,
                             hipStream_t st);
hipError_t launch_aos_to_soa(const void *aos, double *soa, int real_size, size_t n_rays, int nd,
                             hipStream_t st);
hipError_t launch_cast(const double *in, void *out, int real_size, size_t n, hipStream_t st);
hipError_t launch_info_f64(const RayParams &p, const double *data, double *info, size_t n_rays,
                           hipStream_t st);
hipError_t launch_to_bl(double a, double r_match, const double *data, double *out, size_t n_rays,
                        hipStream_t st);
/* (q, k_mu = g(q) u) per ray, AoS [n_rays][8] doubles. */
hipError_t launch_covariant(double a, const double *data, double *out, size_t n_rays,
                            hipStream_t st);

hipError_t pool_f32_blocks_per_cu(int n_freq, int *blocks);
/* Clear n <= 256 Counters with a kernel (no memset node). */
hipError_t launch_zero_counters(Counters *ctr, int n, hipStream_t st);
hipError_t launch_pool_f32(const RayParams &p, const StepRule &sr, double *data, double *info,
                           int64_t *steps, size_t n_rays, const RayRange &rr, Counters *ctr,
                           Fallback *fb, unsigned n_blocks, hipStream_t st);

hipError_t launch_grid_icond(const GridDesc &gd, const RayParams &p, double *data, double *info,
                             int64_t *steps, size_t n_rays, hipStream_t st);
/* Plain FP64 gridded form, n_dump intervals of every ray in one launch
 * (8x8-ray tiles, one per wave, LDS bricks unless bricks = false; the
 * results are the oracle's either way). */
hipError_t launch_grid_tiles_f64(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                                 double *data, double *info, int64_t *steps, size_t n_rays, int W,
                                 int H, bool bricks, hipStream_t st);
hipError_t launch_grid_info_f64(const GridDesc &gd, const double *data, double *info, size_t n_rays,
                                hipStream_t st);
hipError_t launch_grid_covariant(const GridDesc &gd, const double *data, double *out, size_t n_rays,
                                 hipStream_t st);
hipError_t launch_grid_sample_f64(const GridDesc &gd, const double *ev, int which, int comp,
                                  double *out, int n, hipStream_t st);
/* FP32 gridded form, n_dump intervals of every ray in one launch (8x8-ray
 * tiles, one per wave).  bricks = false makes every lane read its stencils
 * from global memory (same values, same order: the results do not change).
 * Steps taken are added to *n_steps. */
hipError_t launch_grid_tiles_f32(const GridDesc &gd, const RtThermal &rt, const StepRule &sr,
                                 double *data, int64_t *steps, size_t n_rays, int W, int H,
                                 bool bricks, unsigned long long *n_steps, hipStream_t st);

} // namespace gray
