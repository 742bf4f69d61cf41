/*
 * serve.h -- device structures of the serving pool.
 *
 * Requests live in a device table.  Each epoch runs three persistent
 * kernels side by side until the epoch's stop time (wall clock), each with
 * its own work queue:
 *   analytic Kerr, FP32 form  one lane per ray; lanes refill from their
 *                             class's queue (rays resumed from the previous
 *                             epoch first, then fresh rays of the class's
 *                             requests) as rays finish;
 *   gridded, FP32 form        one wave per 8x8-ray tile of one request,
 *                             with its LDS brick; waves claim tiles (resumed
 *                             tiles first, then fresh ones);
 *   FP64 re-integration       one lane per ray handed over by the Kerr
 *                             kernel (collapsed step or non-finite state);
 *                             lanes refill from the list.
 * At the stop time a lane or wave saves what it holds, state bits intact,
 * and then carries the resumed entries nobody claimed over unchanged, so a
 * launch may have fewer lanes than it has entries to resume.  Each list is
 * written from the back of its region: the next epoch claims in index
 * order, so the entries written last (those carried over) come first.
 *
 * Kerr requests are grouped in classes that share every parameter except
 * the spin and the camera (channels, step schedule, fluid, escape rule), so
 * that a block of the Kerr kernel serves one class and holds those in scalar
 * registers; spin-derived values are per lane.
 */
#pragma once

#include "grid_device.h"
#include "grayhip_params.h"
#include "rt_thermal.h"
#include "srv_consts.h"

namespace gray {

enum { SRV_NC = 8 + 2 * SRV_NF };     /* reals of a stored (FP64) state          */
enum { SRV_NK = 8 + SRV_NF + 1 };     /* FP32 Kerr state: geodesic, I, shared T  */
enum { SRV_NG = 8 + 2 * SRV_NF };     /* FP32 grid state: geodesic, I/K, tau/K   */

struct SrvRequest {
	RayParams p;          /* physics and camera; nu repeats channel 0 beyond
	                         n_freq                                          */
	StepRule sr;          /* schedule; n_dump intervals                      */
	GridView<float> gv;   /* gridded spacetime (kind == SRV_GRID)            */
	RtThermal rt;         /* gridded transfer constants (nu as p.nu)         */
	int kind, nd, cls, tw;/* nd = 8 + 2 n_freq; Kerr class; tiles per row    */
	double *out_data;     /* [n_rays][nd]                                    */
	double *out_info;     /* [n_rays]                                        */
	int64_t *out_steps;   /* [n_rays]                                        */
};

/* A Kerr class and its share of this epoch's Kerr launch. */
struct SrvClass {
	StepRule sr;
	double nu[SRV_NF];    /* channel frequencies, channel 0 repeated         */
	double n_over_dt;     /* sr.n / sr.dt, sr.dt / sr.n                      */
	double dt_over_n;
	float l2c[SRV_NF];    /* the FP32 transfer's constants (make_rtconst)    */
	float d2[SRV_NF];
	float dts;            /* (float) sr.dts                                  */
	float pad_f;
	double r_esc;
	int n_freq, fluid, escape, pad;
	int block0, block1;   /* blocks [block0, block1) serve this class        */
	int res_off, n_resume;/* resumed rays: in-list [res_off, res_off + n)    */
	int out_base, out_cap;/* its region of the out-list                      */
	int seg0, seg1;       /* fresh segments [seg0, seg1)                     */
	unsigned n_fresh;     /* fresh rays in them                              */
};

/* Rays saved at the end of a Kerr epoch (SoA, capacity cap). */
struct SrvResume {
	int4 *meta;           /* tab, ray, d, h                                  */
	int64_t *steps;
	float *st;            /* [2 SRV_NK][cap]: hi rows, then lo rows          */
	int cap;
};

/* FP64 re-integration entries (SoA, capacity cap). */
struct SrvFb {
	int4 *meta;           /* tab, ray, d, h; h < 0: restart from the camera  */
	int64_t *steps;
	double *x;            /* [SRV_NC][cap]: geodesic, SRV_NF intensities, the
	                         channels' shared depth T (tau_i = d_i T)        */
	int cap;
};

/* Grid tiles saved at the end of an epoch: record r, lane l at r 64 + l. */
struct SrvTiles {
	int2 *hdr;            /* per record: tab, tile                           */
	int4 *brick;          /* per record: the wave's brick origin and cooldown
	                         (reloaded on resume, so that every lane reads its
	                         corners where it would have without the pause) */
	int4 *meta;           /* per lane: ray (-1: none or finished), d, h      */
	int64_t *steps;       /* per lane                                        */
	float *st;            /* [2 SRV_NG][64 cap]                              */
	int cap;              /* records                                         */
};

/* Counters of one epoch (cleared by a kernel before it). */
struct SrvCtr {
	unsigned res_claimed[SRV_MAX_CLASSES];   /* resumed-ray cursor per class  */
	unsigned fresh_claimed[SRV_MAX_CLASSES]; /* fresh-ray cursor per class    */
	unsigned saved[SRV_MAX_CLASSES];         /* rays written to the out region */
	unsigned fb_claimed, fb_saved;      /* FP64 list: cursor; out-list entries
	                                       (survivors, carried, hand-overs) */
	unsigned g_res, g_fresh, g_saved;   /* grid: cursors; records written   */
	unsigned n_done;                    /* requests finished this epoch     */
	unsigned overflow;                  /* writes refused past a region     */
	unsigned heavy_started;             /* FP64 and grid blocks started     */
	unsigned late[3];                   /* waves that started at or after
	                                       t_end: Kerr, FP64, grid          */
	unsigned long long steps[3];        /* FP32 Kerr, FP64 Kerr, grid       */
	unsigned long long t_end;           /* the epoch's stop time (wall clock
	                                       ticks; all ones: none)           */
};

/* Fresh work: segment j covers queue items [start[j], start[j+1]) of its
 * class (Kerr: rays first0 + k of request tab; grid: tiles). */
struct SrvSegs {
	const int *tab;
	const int *first;
	const unsigned *start;
};

struct SrvKerrArgs {
	const SrvRequest *reqs;
	const SrvClass *cls;
	int n_cls;
	SrvSegs seg;
	SrvResume in, out;
	SrvFb fb;             /* the FP64 out-list: hand-overs are written here */
	SrvCtr *ctr;
	unsigned *remaining;  /* per table index: rays not yet finished        */
	int *done;            /* table indices of finished requests             */
	int K;                /* bound on iterations per wave                   */
};

/* The FP64 and grid launches go first; the last of their blocks to start
 * writes `epoch` to *started (host memory), after which the host launches
 * the Kerr kernel into the room they left. */
struct SrvHeavy {
	unsigned total;       /* FP64 + grid blocks of the epoch               */
	unsigned epoch;
	volatile unsigned *started;
};

struct SrvGridArgs {
	const SrvRequest *reqs;
	SrvSegs seg;
	int seg0, seg1;       /* fresh tile segments                           */
	int res_off, n_resume;/* resumed records [res_off, res_off + n_resume)  */
	unsigned n_fresh;     /* fresh tiles                                   */
	int out_cap;          /* records of the out region [0, out_cap)         */
	SrvTiles in, out;
	SrvCtr *ctr;
	unsigned *remaining;
	int *done;
	int K;
	SrvHeavy heavy;
};

struct SrvFbArgs {
	const SrvRequest *reqs;
	SrvFb in;
	int in_off, n_in;     /* entries [in_off, in_off + n_in) of the in-list */
	SrvFb out;            /* survivors and carried entries are written here */
	SrvCtr *ctr;
	unsigned *remaining;
	int *done;
	int K;
	SrvHeavy heavy;
};

} // namespace gray
