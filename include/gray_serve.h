/*
 * gray_serve.h -- the C request loop over the serving pool (libgray.so).
 *
 * Generates a seeded stream of render requests of uneven sizes (the
 * mix: resolutions {64,128,256,512,1024}^2 with weights
 * {.3,.3,.2,.15,.05}, spin U[0, 0.998], cos(inclination) U[0, 1], 1-4
 * frequency channels, 70% analytic Kerr and 30% on one of two resident
 * gridded spacetimes), submits each at its Poisson arrival time, steps the
 * pool until every request is done, and reports per-request latency
 * (arrival to completion) and throughput.  No counterpart in the
 * reference, which renders one image per job.
 */
#ifndef GRAY_SERVE_H
#define GRAY_SERVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	int      n_requests;      /* requests in the stream                        */
	uint64_t seed;            /* stream seed (0x47524159 by default)            */
	double   rate;            /* mean arrivals per second (Poisson)             */
	double   grid_fraction;   /* share of gridded requests (0.3)                */
	int      max_side;        /* cap on the image side (<= 1024), for samples   */
	int      grid_n[2];       /* points per axis of the two grids (96, 128)     */
	size_t   slots;           /* pool size                                     */
	size_t   hbm_budget;      /* bytes of outputs admitted at once             */
	int      epoch_steps;     /* RK4 steps per ray per epoch                   */
	int      device;
	int      verbose;
} gray_serve_opts;

typedef struct {
	int     n_done;
	double  wall_s;           /* first arrival to last completion              */
	double  p50_ms, p99_ms, max_ms, mean_ms;
	int64_t ray_steps[3];     /* FP32 Kerr, FP64 Kerr, FP32 grid               */
	int64_t rays;
	int64_t epochs;
	double  ray_steps_per_s;
} gray_serve_report;

void gray_serve_defaults(gray_serve_opts *o);
/* Run the stream; per-request latencies (ms, in request order) are stored
 * in latency_ms when not NULL (n_requests entries). */
int  gray_serve_run(const gray_serve_opts *o, gray_serve_report *rep, double *latency_ms);

/* One request of the stream, as submitted (for replaying it alone).  The
 * two spacetimes are selected by *grid (0 or 1; -1 for analytic Kerr). */
typedef struct {
	int    side;
	double a_spin, i_obs;
	int    n_freq;
	int    grid;
	double arrival_s;
} gray_serve_item;

int  gray_serve_stream(const gray_serve_opts *o, gray_serve_item *items);

/* Serve a given list of n requests (arrival_s in time order; the first is
 * submitted at once and the others at their offsets from it) through one
 * pool, as gray_serve_run does for the seeded stream.  results (may be
 * NULL, as may any entry) receives request k's rays, AoS [side*side][8 +
 * 2 n_freq] FP64, in results[k]. */
int  gray_serve_run_items(const gray_serve_opts *o, const gray_serve_item *items, int n,
                          gray_serve_report *rep, double *latency_ms, double *const *results);

#ifdef __cplusplus
}
#endif
#endif
