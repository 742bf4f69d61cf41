/*
 * grayhip_params.h -- kernel-side parameter blocks shared by the host layer
 * and the gfx950 kernels.
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/grayhip.h"

namespace gray {

/* Everything a lane needs to know about the render its ray belongs to.
 * Passed by value as a kernel argument (plain render) or read from a
 * request table in HBM (serving pool). */
struct RayParams {
	double a, r_h, eps_stop, r_esc, r_isco;
	int    escape, fluid, n_freq, pad0;
	double nu[GRAYHIP_MAX_FREQ];
	/* camera (sim-org/icond.opts:2-6); trigonometry precomputed on host */
	double w_img, h_img, r_obs, si, ci, sj, cj;
	int    w_rays, h_rays;
};

/* Quantised step schedule of one dump interval (sim-org/driver.cl:81,93). */
struct StepRule {
	double dt;       /* affine length of one interval (dt_dump)        */
	double dts;      /* dt / n_sub: the nominal sub-step                */
	int    n;        /* (INT_MAX / n_sub) * n_sub units per interval    */
	int    n_dump;   /* intervals to run (render); 1 for a single evolve */
};

/* Device-side counters of one FP32 pass (zeroed before it). */
struct Counters {
	unsigned ray_next;                 /* next unclaimed ray of the pool     */
	unsigned n_fb;                     /* entries in the fallback queue      */
	unsigned long long pool_steps;     /* RK4 steps taken by the FP32 pool   */
	unsigned long long fb_steps;       /* RK4 steps taken in FP64 fallback   */
};

/* Rays a pool pass covers: queue index i maps to b0 + i for i < e0 - b0,
 * then to b1 + (i - (e0 - b0)) for the rest, up to e1. */
struct RayRange {
	int b0, e0, b1, e1;
};

/* A ray handed from the FP32 pass to FP64 re-integration. */
struct Fallback {
	int g;           /* ray index                                        */
	int dump;        /* interval it was in                               */
	int h;           /* units of that interval already done              */
	int pad;
};

} // namespace gray
