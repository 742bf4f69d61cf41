/*
 * grayhip.h -- C ABI of the MI355X (gfx950) GRay2 device layer.
 *
 * This is the drop-in boundary below the job module: every entry point
 * replaces one device-side service that the reference's host code obtains
 * from lux-OpenCL.  Plain pointers and sizes only; all functions return 0 on
 * success and a negative GRAYHIP_E* code on failure, with a message kept per
 * context (grayhip_last_error) or per thread (grayhip_global_error).
 *
 * Ray order and pixel mapping are those of the reference: ray g = gi + gj*W
 * for column gi < W and row gj < H (sim-org/driver.cl:47-57).  The state of a
 * ray is n_data = 8 + 2*n_freq reals  t x y z u^t u^x u^y u^z I[n_freq]
 * tau[n_freq]  (sim-org/phys.cl:30-33, rt.cl:174-177); info holds g(u,u).
 */
#ifndef GRAYHIP_H
#define GRAYHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRAYHIP_MAX_FREQ 8

enum {
	GRAYHIP_OK        = 0,
	GRAYHIP_EINVAL    = -1,  /* bad argument or configuration          */
	GRAYHIP_ENOMEM    = -2,  /* device allocation failed               */
	GRAYHIP_EDEVICE   = -3,  /* HIP runtime error                      */
	GRAYHIP_ESTATE    = -4,  /* call out of order (e.g. evolve before icond) */
	GRAYHIP_ENODEV    = -5   /* no usable GPU                          */
};

/* Arithmetic forms (the `precision` option, sim-org/setup.opts:6). */
enum {
	GRAYHIP_FP32 = 4,  /* FP32 bulk RK4/RHS, compensated state, FP64 ICs,
	                      FP64 re-integration of collapsed-step rays      */
	GRAYHIP_FP64 = 8   /* plain FP64 form, bitwise equal to the oracle    */
};

/* Termination (sim-org/phys.cl:48-64). */
enum {
	GRAYHIP_ESC_REF     = 0,  /* as written: eps < eps_stop only (the escape
	                             test of phys.cl:57 can never fire)       */
	GRAYHIP_ESC_OUTWARD = 1   /* plus r > r_esc while moving outward along
	                             the integration direction                */
};

/* Fluid and emission model feeding radiative transfer. */
enum {
	GRAYHIP_FLUID_NONE    = 0,
	GRAYHIP_FLUID_DISK    = 1,  /* synthetic Keplerian disk, j = n nu^-1/2,
	                               alpha = 0.1 n nu^-5/2 (analytic Kerr)   */
	GRAYHIP_FLUID_UNIFORM = 2,  /* n = 1, shift = 1 (slab closed forms)     */
	GRAYHIP_FLUID_THERMAL = 3   /* the reference model, sim-org/flow.cl:36-58
	                               with rt.cl:80-198: n_e from the
	                               spacetime's density, T_e = 1e12, fluid
	                               at rest, field b^mu = (0,0,0,b_field)
	                               (b = 0 in the reference); bremsstrahlung
	                               and synchrotron emission (the latter
	                               vanishes at b = 0), Kirchhoff
	                               absorption.  Gridded only.              */
};

/* Grid coordinate maps of a numerical spacetime (sim-org/interp.cl:69-121
 * hard-codes the fisheye one; uniform axes are the other case in use). */
enum {
	GRAYHIP_MAP_UNIFORM = 0,  /* x_i = lo + i (hi - lo)/(n - 1)           */
	GRAYHIP_MAP_FISHEYE = 1   /* x_i = sinh((A i + B)^3), A, B from lo, hi */
};

/* A numerical spacetime resident on one GPU: two time-level slots of
 * 40 Christoffels + 10 metric components per grid point. */
typedef struct grayhip_spacetime grayhip_spacetime;

typedef struct {
	int    n[3];             /* points along x, y, z                         */
	double lo[3], hi[3];     /* first and last axis values (the bounding box,
	                            sim-org/io.c:266-274)                        */
	int    map;              /* GRAYHIP_MAP_*                                */
	int    dtype;            /* bytes per stored real: 4 or 8                */
	int    device;
} grayhip_grid_desc;

typedef struct {
	size_t w_rays, h_rays;   /* W, H (sim-org/param.opts:10-11)            */
	int    precision;        /* GRAYHIP_FP64 or GRAYHIP_FP32               */
	double a_spin;           /* sim/Kerr.opts:2                            */
	int    n_freq;           /* 0 turns radiative transfer off             */
	double nu[GRAYHIP_MAX_FREQ];
	int    fluid;            /* GRAYHIP_FLUID_*                            */
	int    escape;           /* GRAYHIP_ESC_*                              */
	double r_esc;            /* escape radius for GRAYHIP_ESC_OUTWARD      */
	double eps_stop;         /* horizon stop r - r_H < eps_stop; <= 0 picks
	                            0.01 r_H for ESC_OUTWARD, 0.01 for ESC_REF  */
	int    device;           /* HIP device ordinal                          */
	/* numerical spacetime (coordinates=dyst); NULL: analytic Kerr-Schild.
	 * With a spacetime, transfer needs GRAYHIP_FLUID_THERMAL and uploaded
	 * densities; `escape` selects box-exit termination with the horizon
	 * proxy (ESC_OUTWARD) or none (ESC_REF, as sim-org/dyst.cl). */
	grayhip_spacetime *spacetime;
	double t_init;           /* grid time of the rays' q^t = 0              */
	int    raw_icond;        /* 1: no null normalisation (dyst.cl:89-110)   */
	double m_adm;            /* M_ADM in solar masses for FLUID_THERMAL
	                            (sim-org/param.opts:4); <= 0 picks 4.3e6    */
	double b_field;          /* FLUID_THERMAL's field b^z in Gauss: 0 is the
	                            reference flow (flow.cl:47, b = 0); > 0 adds
	                            the synchrotron term of rt.cl:145-170     */
} grayhip_config;

typedef struct grayhip_ctx grayhip_ctx;

/* Allocate device state for W*H rays (replaces the buffer set-up of
 * sim-org/gray.c:130-133 and the kernel specialisation of build.c:30-143:
 * resolution, spin and channels are run-time arguments here). */
int  grayhip_create(const grayhip_config *cfg, grayhip_ctx **out);

/* Free everything (sim-org/gray.c:330-346). */
void grayhip_destroy(grayhip_ctx *ctx);

/* Camera initial conditions in FP64 (replaces sim-org/icond.c:29-86 and
 * the icond_drv kernel, driver.cl:36-67).  Angles in degrees. */
int  grayhip_icond(grayhip_ctx *ctx, double w_img, double h_img,
                   double r_obs, double i_obs, double j_obs);

/* One dump interval: every ray advances by affine parameter dt in the
 * integer-quantised sub-steps of n_sub nominal steps (replaces
 * sim-org/evolve.c:28-67 and evolve_drv, driver.cl:70-108).  *ns (may be
 * NULL) receives the device time in nanoseconds, as the reference's
 * evolve() returns it. */
int  grayhip_evolve(grayhip_ctx *ctx, double dt, size_t n_sub, double *ns);

/* The whole job after icond: n_dump intervals of dt each, integrated in one
 * persistent ray-pool pass (rays retire and lanes refill independently).
 * Equivalent to n_dump calls of grayhip_evolve without the intermediate
 * dumps.  *ns receives the device time. */
int  grayhip_render(grayhip_ctx *ctx, double dt, size_t n_sub, size_t n_dump,
                    double *ns);

/* Copy the ray states to the host in the reference's dump order
 * (AoS [H][W][n_data], sim-org/io.c:61-66) as `precision`-byte reals;
 * info (may be NULL) receives n_rays reals of g(u,u)
 * (replaces the ocl->mmap readback of io.c:52-53). */
int  grayhip_read(grayhip_ctx *ctx, void *data, void *info);

/* Same as grayhip_read but always in FP64 (the device keeps FP64 state in
 * both forms): data AoS [H][W][n_data] doubles, info n_rays doubles. */
int  grayhip_read_f64(grayhip_ctx *ctx, double *data, double *info);

/* Boyer-Lindquist view of the states (analytic Kerr only; replaces the
 * getBL/getBLu diagnostics of sim-org/KS.cl:63-149, which need the
 * undefined r_match): AoS [H][W][8] doubles (t, r, theta, phi, u^t, u^r,
 * u^theta, u^phi) with t matched at r = r_match (> r_H) and phi at
 * infinity.  Spheroidal convention x + i y = (r + i a) e^{i phi} sin(theta),
 * the one in which the metric maps to BL by r-only shifts; the velocity
 * uses u^r in the t and phi corrections.  Rays inside the horizon give NaN. */
int  grayhip_read_bl(grayhip_ctx *ctx, double r_match, double *out);

/* Covariant view of the states: AoS [H][W][8] doubles (t, x, y, z, k_t,
 * k_x, k_y, k_z), k_mu = g_{mu nu}(q) u^nu, the output the reference
 * obtains with down() (sim-org/KS.cl:91-114 on analytic Kerr-Schild,
 * dyst.cl:55-81 with the interpolated metric on a numerical spacetime);
 * computed in FP64 from the stored state in either form. */
int  grayhip_read_covariant(grayhip_ctx *ctx, double *out);

/* Upload AoS states (same layout as grayhip_read) to restart from a dump.
 * With the disk or uniform fluid every channel's optical depth is one
 * shared integral T (tau_i = 0.1 nu_i^-5/2 T, as those fluids absorb): a
 * state whose depths are not in that proportion (to 64 rounding errors of
 * the real size) is rejected with GRAYHIP_EINVAL. */
int  grayhip_write(grayhip_ctx *ctx, const void *data);

/* Per-ray count of RK4 steps executed since icond (int64[n_rays]). */
int  grayhip_read_steps(grayhip_ctx *ctx, int64_t *steps);

/* Counters of the last render/evolve: total RK4 steps, and the number of
 * rays re-integrated in FP64 (FP32 form only). */
int  grayhip_stats(grayhip_ctx *ctx, int64_t *steps_total,
                   int64_t *fallback_rays);

/* RK4 steps of the last FP32-form call split by where they ran: the FP32
 * pool pass and the FP64 re-integration (both 0 after an FP64-form call). */
int  grayhip_pass_steps(grayhip_ctx *ctx, int64_t *pool_steps,
                        int64_t *fallback_steps);

/* Device time (ns) of the dominant kernel in the last call, measured with
 * HIP events on the stream it ran on; *launches receives its launch count. */
int  grayhip_kernel_time(grayhip_ctx *ctx, double *ns, int *launches);

/* ---- numerical spacetimes (replace the 2 x 51 image3d_t objects of
 * sim-org/gray.h:92-93 and their loaders, io.c:302-533) ---- */

/* Allocate both level slots on the device. */
int  grayhip_spacetime_create(const grayhip_grid_desc *d, grayhip_spacetime **out);
void grayhip_spacetime_destroy(grayhip_spacetime *st);
/* Copy one time level (host arrays of d->dtype reals, interleaved per grid
 * point p = (k*ny + j)*nx + i: gam[p*40 + 10*mu + c] with c in
 * tt tx ty tz xx xy xz yy yz zz, met[p*10 + c] in the same order) into slot
 * 0 (t1) or 1 (t2) at the given time.  Blocks until the copy is done. */
int  grayhip_spacetime_upload(grayhip_spacetime *st, int slot, double time,
                              const void *gam, const void *met);
/* Copy the density (the rho dataset, n_points reals of d->dtype in the
 * same point order) of the level held in `slot`; GRAYHIP_FLUID_THERMAL
 * reads it as n_e (sim-org/flow.cl:50). */
int  grayhip_spacetime_upload_rho(grayhip_spacetime *st, int slot, const void *rho);
/* Slide the window: to_t1 != 0 makes the t2 slot the new t1 slot, else the
 * t1 slot the new t2 slot (sim-org/io.c:279-300); the other slot is then
 * free for the next upload. */
int  grayhip_spacetime_shift(grayhip_spacetime *st, int to_t1);
/* frozen != 0: space-only interpolation from slot 0 (sim-org/gray.c:210-228). */
int  grayhip_spacetime_set_frozen(grayhip_spacetime *st, int frozen);
/* Device bytes held by the spacetime. */
size_t grayhip_spacetime_bytes(const grayhip_spacetime *st);
/* Interpolate one field at n events (t,x,y,z) on the device in FP64:
 * which 0: Christoffel component comp < 40, 1: metric comp < 10,
 * 2: density (comp 0). */
int  grayhip_spacetime_sample(grayhip_spacetime *st, int n, const double *events,
                              int which, int comp, double t_init, double *out);
const char *grayhip_spacetime_error(const grayhip_spacetime *st);

/* ---- serving: many renders of uneven size share one ray pool ----
 *
 * No counterpart in the reference (one job renders one image).  Requests
 * are queued by grayhip_server_submit and admitted when their output fits
 * in the HBM budget; each grayhip_server_step runs one epoch: finished rays
 * free their slots, freed slots are refilled with rays of admitted
 * requests (initial conditions in FP64 on the device), and every occupied
 * slot advances by at most `epoch_steps` RK4 steps.  Analytic-Kerr rays run
 * in the FP32 form and move to FP64 re-integration when their step
 * collapses; gridded rays run in the FP32 form on resident spacetimes.  A
 * ray's arithmetic does not depend on which slot, epoch or neighbours it
 * gets, so a request's result is the same alone and in any mix.
 *
 * Threads: one thread calls grayhip_server_step; submit, next_done, state,
 * result and stats may be called from other threads at the same time (the
 * host bookkeeping is locked, and the lock is not held while an epoch runs
 * on the GPU).  create and destroy are not concurrent with anything. */
typedef struct grayhip_server grayhip_server;

typedef struct {
	int    device;
	size_t slots;            /* rays in flight (pool size)                  */
	size_t hbm_budget;       /* bytes of request outputs admitted at once   */
	int    epoch_steps;      /* RK4 steps per ray per epoch                 */
	int    max_requests;     /* admitted requests at once                   */
} grayhip_server_config;

typedef struct {
	size_t w_rays, h_rays;
	double a_spin;           /* analytic Kerr (ignored with a spacetime)     */
	int    n_freq;           /* 0..4                                         */
	double nu[4];
	int    fluid, escape;
	double r_esc, eps_stop;
	double w_img, h_img, r_obs, i_obs, j_obs;
	double dt;               /* affine length of one dump interval           */
	size_t n_sub, n_dump;
	grayhip_spacetime *spacetime; /* FP32 storage; NULL: analytic Kerr       */
	double t_init;
} grayhip_request;

enum { GRAYHIP_REQ_QUEUED = 0, GRAYHIP_REQ_RUNNING = 1, GRAYHIP_REQ_DONE = 2 };

int  grayhip_server_create(const grayhip_server_config *cfg, grayhip_server **out);
void grayhip_server_destroy(grayhip_server *srv);
/* Queue a request; *id identifies it.  Output needs
 * n_rays * ((8 + 2 n_freq) + 2) * 8 bytes of the budget. */
int  grayhip_server_submit(grayhip_server *srv, const grayhip_request *req, int64_t *id);
/* One epoch (blocks until it is done).  *busy is set to 0 when no request
 * is queued or running. */
int  grayhip_server_step(grayhip_server *srv, int *busy);
/* Next finished request not yet reported (returns 0 and *id, or 1 if none). */
int  grayhip_server_next_done(grayhip_server *srv, int64_t *id);
int  grayhip_server_state(grayhip_server *srv, int64_t id, int *state);
/* Copy a finished request's rays (AoS [n_rays][8 + 2 n_freq] FP64, info,
 * RK4 steps per ray; any may be NULL) and release it. */
int  grayhip_server_result(grayhip_server *srv, int64_t id, double *data, double *info,
                           int64_t *steps);
/* Totals since creation: RK4 steps (FP32 Kerr, FP64 Kerr, gridded), epochs. */
int  grayhip_server_stats(grayhip_server *srv, int64_t steps[3], int64_t *epochs);
/* Message of the calling thread's last failed grayhip_server_* call. */
const char *grayhip_server_error(const grayhip_server *srv);

/* ---- a render that allocates nothing ----
 *
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  icond + render of grayhip_icond / grayhip_render for
 * an analytic-Kerr config (spacetime must be NULL) into caller-owned device
 * buffers: data SoA [8 + 2 n_freq][W*H] doubles (state s of ray g at
 * data[s*W*H + g]), info [W*H] doubles of g(u,u), steps [W*H] RK4 steps per
 * ray.  The FP32 form's scratch (pass counters and the FP64 re-integration
 * queue, 16 B per ray) is a fixed device array inside the library, one per
 * GPU, held until the process ends: W*H may not exceed
 * GRAYHIP_STATIC_MAX_RAYS (GRAYHIP_EINVAL otherwise).  Work is queued on
 * `stream` (a hipStream_t; NULL is the null stream) and the call returns
 * without waiting; results equal grayhip_render's bit for bit.  Calls that
 * overlap on one GPU share that scratch and race: serialise them. */
#define GRAYHIP_STATIC_MAX_RAYS (1 << 20)
int  grayhip_render_static(const grayhip_config *cfg, double w_img, double h_img,
                           double r_obs, double i_obs, double j_obs, double dt,
                           size_t n_sub, size_t n_dump, double *data, double *info,
                           int64_t *steps, void *stream);

int  grayhip_n_data(const grayhip_ctx *ctx);
const char *grayhip_last_error(const grayhip_ctx *ctx);
const char *grayhip_global_error(void);

/* Number of visible HIP devices (0 when none); never aborts. */
int  grayhip_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
