/*
 * serve.h -- device structures of the serving pool: the request table, the
 * slot arrays (one ray per slot, SoA) and the per-epoch output lists.
 */
#pragma once

#include "grid_device.h"
#include "grayhip_params.h"

namespace gray {

enum { SRV_KERR = 0, SRV_GRID = 1 };
enum { SRV_NF = 4 };                  /* transfer channels carried per Kerr slot */
enum { SRV_NC = 8 + 2 * SRV_NF };     /* components of a Kerr slot state         */

struct SrvRequest {
	RayParams p;          /* physics and camera (n_freq <= SRV_NF)          */
	StepRule sr;          /* per-request schedule; n_dump intervals         */
	GridView<float> gv;   /* gridded spacetime (kind == SRV_GRID)           */
	int kind;
	int nd;               /* 8 + 2 n_freq: components in the output         */
	int pad[2];
	double *out_data;     /* [n_rays][nd]                                   */
	double *out_info;     /* [n_rays]                                       */
	int64_t *out_steps;   /* [n_rays]                                       */
};

/* Slot arrays of capacity S.  st holds hi (rows 0..SRV_NC-1) and lo (rows
 * SRV_NC..2 SRV_NC-1) of the FP32 state, st64 the FP64 state of slots in
 * re-integration; h < 0 there asks for a restart from the camera. */
struct SrvSlots {
	int *req, *ray, *d, *h;
	int64_t *steps;
	float *st;
	double *st64;
	int S;
};

struct SrvCounters {
	unsigned n_fin;                  /* slots finished this epoch            */
	unsigned n_to64;                 /* slots moved to FP64 this epoch        */
	unsigned long long steps[3];     /* RK4 steps: FP32 Kerr, FP64 Kerr, grid */
};

struct SrvLists {
	int *fin;                        /* finished slot indices                */
	int *to64;                       /* slots moved to FP64                  */
	SrvCounters *ctr;
};

/* Per-lane parameters of the FP32 Kerr path (the fields getdt and the
 * transfer RHS use). */
struct LaneKerr {
	float a, r_h, eps_stop, r_esc, r_isco;
	int escape, fluid, n_freq;
	float nu[SRV_NF];
};

} // namespace gray
