/*
 * grayhip_host.hip -- the C-ABI device layer (include/grayhip.h).
 *
 * Owns the device buffers of one render (SoA FP64 ray state, info, per-ray
 * step counters, the FP32 pass's fallback queue) and one HIP stream, and
 * sequences the kernels.  Replaces the lux-OpenCL services used by
 * sim-org/gray.c, icond.c, evolve.c and io.c:dump.
 */
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "launch.h"
#include "spacetime.h"

using namespace gray;

/* Pool passes of the FP32 form over bands of image rows (run_fp32). */
constexpr int FP32_PASSES = 3;

struct grayhip_ctx {
	grayhip_config cfg;
	RayParams p;
	size_t n_rays;
	int nd;
	double *d_data;
	double *d_info;
	int64_t *d_steps;
	Counters *d_ctr;      /* counters of the FP32 pass                       */
	Fallback *d_fb;
	void *d_tmp;          /* AoS staging for read/write                      */
	hipStream_t stream, stream2, stream3;
	hipEvent_t ev[12];
	bool have_icond;
	int pool_blocks;
	grayhip_spacetime *st;  /* numerical spacetime, or NULL (analytic Kerr) */
	RtThermal rt;           /* emission model constants (gridded transfer)  */
	int64_t last_fallback;
	Counters last_ctr;
	double kern_ns;
	int kern_launches;
	char err[512];
};

static thread_local char g_err[512];

static int fail(grayhip_ctx *c, int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(c ? c->err : g_err, 512, fmt, ap);
	va_end(ap);
	return code;
}

#define HCK(ctx, expr)                                                                    \
	do {                                                                              \
		hipError_t e_ = (expr);                                                   \
		if (e_ != hipSuccess)                                                     \
			return fail(ctx, GRAYHIP_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
	} while (0)

static GridDesc grid_view(const grayhip_ctx *c);

/* Constants of sim-org/rt.cl:28-143 in cgs (CODATA 2010 values, as there),
 * evaluated with the same expressions as the CPU oracle. */
void gray::rt_thermal_init(RtThermal *r, double m_adm, double b, int n_freq, const double *nu)
{
	const double c_light = 2.99792458e+10, h = 6.62606957e-27, G = 6.67384800e-08;
	const double k_b = 1.38064881e-16, ry = 2.17987197e-11, m_e = 9.10938291e-28;
	const double q_e = 4.80320425e-10;
	const double m_sun = 1.98910000e+33, zeta = 1.78107241799;
	const double pi = 3.14159265358979323846;
	memset(r, 0, sizeof(*r));
	r->planck1 = (2.0 * h) * c_light;
	r->planck2 = h / (m_e * c_light);
	r->c = c_light;
	r->gaunt_a = sqrt(3.0 / pi);
	r->gaunt_b = sqrt(3.0) / pi;
	r->gaunt_c = log(4.0 / zeta);
	r->gaunt_d = log(4.0 / pow(zeta, 2.5));
	r->gaunt_e = sqrt(12.0);
	r->ff_x = ((m_e * c_light) * c_light) / ry;
	r->ff_y = h / ((m_e * c_light) * c_light);
	r->ff_f = sqrt(((G * m_sun) / (c_light * c_light)) * 6.8e-38 /
	               ((4.0 * pi) * sqrt(((m_e * c_light) * c_light) / k_b)));
	r->m_adm = m_adm;
	r->te = RT_TE_FLOW;
	r->scale = m_adm * (r->ff_f * r->ff_f);
	r->n_freq = n_freq;
	for (int i = 0; i < n_freq; ++i)
		r->nu[i] = nu[i];
	/* synchrotron, rt.cl:145-170; T_e lies above the ln K_2 table, where
	 * ln K_2 = ln(2 T_e^2 - 1/2) */
	r->b = b;
	r->te2b = (r->te * r->te) * b;
	r->syn_nus = q_e / (((9.0 * pi) * m_e) * c_light);
	r->syn_f = ((G * m_sun) / (c_light * c_light)) *
	           ((sqrt(2.0) * pi) * (q_e * q_e) / (3.0 * c_light));
	r->syn_k = pow(2.0, 11.0 / 12.0);
	r->e_lk2 = exp(-log((2.0 * r->te) * r->te - 0.5));
}

/* Prograde ISCO (Bardeen, Press & Teukolsky 1972), M = 1. */
static double r_isco(double a)
{
	double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
	double z2 = sqrt(3.0 * a * a + z1 * z1);
	return 3.0 + z2 - sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

extern "C" int grayhip_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess)
		return 0;
	return n;
}

extern "C" const char *grayhip_global_error(void) { return g_err; }

extern "C" const char *grayhip_last_error(const grayhip_ctx *ctx)
{
	return ctx ? ctx->err : g_err;
}

extern "C" int grayhip_n_data(const grayhip_ctx *ctx) { return ctx ? ctx->nd : -1; }

static void release(grayhip_ctx *c)
{
	if (!c)
		return;
	(void)hipFree(c->d_data);
	(void)hipFree(c->d_info);
	(void)hipFree(c->d_steps);
	(void)hipFree(c->d_ctr);
	(void)hipFree(c->d_fb);
	(void)hipFree(c->d_tmp);
	for (int i = 0; i < 12; ++i)
		if (c->ev[i])
			(void)hipEventDestroy(c->ev[i]);
	if (c->stream)
		(void)hipStreamDestroy(c->stream);
	if (c->stream2)
		(void)hipStreamDestroy(c->stream2);
	if (c->stream3)
		(void)hipStreamDestroy(c->stream3);
	free(c);
}

extern "C" void grayhip_destroy(grayhip_ctx *ctx)
{
	if (ctx)
		(void)hipSetDevice(ctx->cfg.device);
	release(ctx);
}

/* Arguments of a render, checked before anything is allocated (messages
 * go to the calling thread's error string). */
static int check_config(const grayhip_config *cfg)
{
	if (cfg->w_rays == 0 || cfg->h_rays == 0)
		return fail(nullptr, GRAYHIP_EINVAL, "w_rays and h_rays must be positive");
	if (cfg->w_rays > INT_MAX || cfg->h_rays > INT_MAX ||
	    cfg->w_rays * cfg->h_rays > (size_t)INT_MAX)
		return fail(nullptr, GRAYHIP_EINVAL, "too many rays (W*H must fit in int)");
	if (cfg->precision != GRAYHIP_FP32 && cfg->precision != GRAYHIP_FP64)
		return fail(nullptr, GRAYHIP_EINVAL, "precision must be 4 or 8, got %d", cfg->precision);
	if (cfg->n_freq < 0 || cfg->n_freq > GRAYHIP_MAX_FREQ)
		return fail(nullptr, GRAYHIP_EINVAL, "n_freq must be in [0, %d]", GRAYHIP_MAX_FREQ);
	if (!(fabs(cfg->a_spin) < 1.0))
		return fail(nullptr, GRAYHIP_EINVAL, "|a_spin| must be < 1");
	if (cfg->escape != GRAYHIP_ESC_REF && cfg->escape != GRAYHIP_ESC_OUTWARD)
		return fail(nullptr, GRAYHIP_EINVAL, "unknown escape rule %d", cfg->escape);
	if (cfg->fluid < GRAYHIP_FLUID_NONE || cfg->fluid > GRAYHIP_FLUID_THERMAL)
		return fail(nullptr, GRAYHIP_EINVAL, "unknown fluid model %d", cfg->fluid);
	if (cfg->spacetime && cfg->n_freq > 0 && cfg->fluid != GRAYHIP_FLUID_THERMAL)
		return fail(nullptr, GRAYHIP_EINVAL, "radiative transfer on a numerical spacetime uses "
		                                     "the thermal fluid (the reference flow)");
	if (!cfg->spacetime && cfg->n_freq > 0 && cfg->fluid == GRAYHIP_FLUID_THERMAL)
		return fail(nullptr, GRAYHIP_EINVAL, "the thermal fluid reads the density of a "
		                                     "numerical spacetime");
	if (cfg->spacetime && cfg->spacetime->desc.device != cfg->device)
		return fail(nullptr, GRAYHIP_EINVAL, "spacetime lives on another device");
	if (cfg->n_freq > 0 && cfg->fluid == GRAYHIP_FLUID_NONE)
		return fail(nullptr, GRAYHIP_EINVAL, "radiative transfer needs a fluid model");
	for (int i = 0; i < cfg->n_freq; ++i)
		if (!(cfg->nu[i] > 0.0))
			return fail(nullptr, GRAYHIP_EINVAL, "frequency channel %d must be positive", i);
	if (!(cfg->b_field >= 0.0 && cfg->b_field < 1e30))
		return fail(nullptr, GRAYHIP_EINVAL, "b_field must be finite and >= 0");
	if (cfg->b_field > 0.0 && !(cfg->fluid == GRAYHIP_FLUID_THERMAL && cfg->n_freq > 0))
		return fail(nullptr, GRAYHIP_EINVAL, "b_field belongs to the thermal fluid's transfer "
		                                     "(gridded spacetimes with channels)");

	int ndev = grayhip_device_count();
	if (ndev <= 0)
		return fail(nullptr, GRAYHIP_ENODEV, "no HIP device visible");
	if (cfg->device < 0 || cfg->device >= ndev)
		return fail(nullptr, GRAYHIP_EINVAL, "device %d out of range (%d visible)", cfg->device, ndev);

	return GRAYHIP_OK;
}

/* Physics and image size of a render (the camera comes with icond). */
static void fill_params(const grayhip_config *cfg, RayParams &p)
{
	memset(&p, 0, sizeof(p));
	p.a = cfg->a_spin;
	p.r_h = 1.0 + sqrt(1.0 - p.a * p.a);
	p.escape = cfg->escape;
	p.eps_stop = cfg->eps_stop > 0.0 ? cfg->eps_stop
	             : (cfg->escape == GRAYHIP_ESC_OUTWARD ? 0.01 * p.r_h : 0.01);
	p.r_esc = cfg->r_esc;
	p.r_isco = r_isco(p.a);
	p.fluid = cfg->n_freq > 0 ? cfg->fluid : GRAYHIP_FLUID_NONE;
	p.n_freq = cfg->n_freq;
	for (int i = 0; i < cfg->n_freq; ++i)
		p.nu[i] = cfg->nu[i];
	p.w_rays = (int)cfg->w_rays;
	p.h_rays = (int)cfg->h_rays;
}

/* Camera of sim-org/icond.opts:2-6, trigonometry in FP64 on the host. */
static void fill_camera(RayParams &p, double w_img, double h_img, double r_obs, double i_obs,
                        double j_obs)
{
	const double d2r = M_PI / 180.0;
	p.w_img = w_img;
	p.h_img = h_img;
	p.r_obs = r_obs;
	p.si = sin(d2r * i_obs);
	p.ci = cos(d2r * i_obs);
	p.sj = sin(d2r * j_obs);
	p.cj = cos(d2r * j_obs);
}

extern "C" int grayhip_create(const grayhip_config *cfg, grayhip_ctx **out)
{
	if (!cfg || !out)
		return fail(nullptr, GRAYHIP_EINVAL, "null argument");
	*out = nullptr;
	{
		int rc0 = check_config(cfg);
		if (rc0)
			return rc0;
	}

	grayhip_ctx *c = (grayhip_ctx *)calloc(1, sizeof(grayhip_ctx));
	if (!c)
		return fail(nullptr, GRAYHIP_ENOMEM, "host allocation failed");
	c->cfg = *cfg;
	c->n_rays = cfg->w_rays * cfg->h_rays;
	c->nd = 8 + 2 * cfg->n_freq;

	fill_params(cfg, c->p);
	c->st = cfg->spacetime;
	rt_thermal_init(&c->rt, cfg->m_adm > 0.0 ? cfg->m_adm : 4.3e6, cfg->b_field,
	                c->st ? cfg->n_freq : 0, cfg->nu);

	int rc;
	hipError_t e;
	if ((e = hipSetDevice(cfg->device)) != hipSuccess) {
		rc = fail(nullptr, GRAYHIP_EDEVICE, "hipSetDevice: %s", hipGetErrorString(e));
		goto bad;
	}
	{
		size_t n = c->n_rays, nd = (size_t)c->nd;
		struct { void **ptr; size_t bytes; } al[] = {
			{(void **)&c->d_data, n * nd * sizeof(double)},
			{(void **)&c->d_info, n * sizeof(double)},
			{(void **)&c->d_steps, n * sizeof(int64_t)},
			{(void **)&c->d_ctr, FP32_PASSES * sizeof(Counters)},
			{(void **)&c->d_fb, n * sizeof(Fallback)},
			{(void **)&c->d_tmp, n * nd * sizeof(double)},
		};
		for (auto &a : al) {
			if ((e = hipMalloc(a.ptr, a.bytes)) != hipSuccess) {
				rc = fail(nullptr, GRAYHIP_ENOMEM, "hipMalloc(%zu): %s", a.bytes,
				          hipGetErrorString(e));
				goto bad;
			}
		}
	}
	if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
	    (e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking)) != hipSuccess ||
	    (e = hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking)) != hipSuccess) {
		rc = fail(nullptr, GRAYHIP_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
		goto bad;
	}
	for (int i = 0; i < 12; ++i)
		if ((e = hipEventCreate(&c->ev[i])) != hipSuccess) {
			rc = fail(nullptr, GRAYHIP_EDEVICE, "hipEventCreate: %s", hipGetErrorString(e));
			goto bad;
		}
	{
		int bpc = 0, n_cu = 0;
		if ((e = pool_f32_blocks_per_cu(c->p.n_freq, &bpc)) != hipSuccess ||
		    (e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount,
		                               cfg->device)) != hipSuccess) {
			rc = fail(nullptr, GRAYHIP_EDEVICE, "occupancy query: %s", hipGetErrorString(e));
			goto bad;
		}
		if (bpc < 1)
			bpc = 1;
		size_t need = (c->n_rays + 255) / 256;
		size_t full = (size_t)bpc * (size_t)n_cu;
		c->pool_blocks = (int)(need < full ? need : full);
	}
	*out = c;
	return GRAYHIP_OK;
bad:
	release(c);
	return rc;
}

extern "C" int grayhip_icond(grayhip_ctx *c, double w_img, double h_img, double r_obs,
                             double i_obs, double j_obs)
{
	if (!c)
		return fail(nullptr, GRAYHIP_EINVAL, "null context");
	if (!(r_obs > 0.0))
		return fail(c, GRAYHIP_EINVAL, "r_obs must be positive");
	fill_camera(c->p, w_img, h_img, r_obs, i_obs, j_obs);
	HCK(c, hipSetDevice(c->cfg.device));
	if (c->st && c->rt.n_freq > 0 && !grayhip_spacetime_has_rho(c->st))
		return fail(c, GRAYHIP_ESTATE, "transfer needs the density of every level in use "
		                               "(grayhip_spacetime_upload_rho)");
	if (c->st)
		HCK(c, launch_grid_icond(grid_view(c), c->p, c->d_data, c->d_info, c->d_steps, c->n_rays,
		                         c->stream));
	else
		HCK(c, launch_icond(c->p, c->d_data, c->d_info, c->d_steps, c->n_rays, c->stream));
	HCK(c, hipStreamSynchronize(c->stream));
	c->have_icond = true;
	return GRAYHIP_OK;
}

static GridDesc grid_view(const grayhip_ctx *c)
{
	return spacetime_view(c->st, c->cfg.t_init, c->cfg.raw_icond,
	                      c->cfg.escape == GRAYHIP_ESC_REF);
}

/* Numerical spacetime: both forms run every interval of the call in one
 * launch of their tile kernel (LDS-staged stencils); the FP32 form's
 * diagnostics are recomputed in FP64 at the end. */
static int run_grid(grayhip_ctx *c, const StepRule &one, size_t n_dump, double *ns)
{
	if (c->rt.n_freq > 0 && !grayhip_spacetime_has_rho(c->st))
		return fail(c, GRAYHIP_ESTATE, "transfer needs the density of every level in use "
		                               "(grayhip_spacetime_upload_rho)");
	const GridDesc g = grid_view(c);
	/* GRAYHIP_GRID_BRICKS=0 turns the LDS bricks off (a diagnostic: the
	 * results are the same either way) */
	const char *env_b = getenv("GRAYHIP_GRID_BRICKS");
	const bool bricks = !(env_b && env_b[0] == '0');
	HCK(c, hipMemsetAsync(c->d_ctr, 0, sizeof(Counters), c->stream));
	HCK(c, hipEventRecord(c->ev[0], c->stream));
	/* all intervals in one launch: one front of rays, no per-interval tail
	 * (the FP64 state stays in registers: the same bits as one launch per
	 * interval) */
	StepRule all = one;
	all.n_dump = (int)n_dump;
	if (c->cfg.precision == GRAYHIP_FP64) {
		HCK(c, launch_grid_tiles_f64(g, c->rt, all, c->d_data, c->d_info, c->d_steps, c->n_rays,
		                             (int)c->cfg.w_rays, (int)c->cfg.h_rays, bricks, c->stream));
	} else {
		/* no FP64 round trip of the compensated state between intervals */
		HCK(c, launch_grid_tiles_f32(g, c->rt, all, c->d_data, c->d_steps, c->n_rays,
		                             (int)c->cfg.w_rays, (int)c->cfg.h_rays, bricks,
		                             &c->d_ctr->pool_steps, c->stream));
	}
	HCK(c, hipEventRecord(c->ev[1], c->stream));
	if (c->cfg.precision != GRAYHIP_FP64)
		HCK(c, launch_grid_info_f64(g, c->d_data, c->d_info, c->n_rays, c->stream));
	HCK(c, hipMemcpyAsync(&c->last_ctr, c->d_ctr, sizeof(Counters), hipMemcpyDeviceToHost,
	                      c->stream));
	HCK(c, hipStreamSynchronize(c->stream));
	float ms = 0.f;
	HCK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
	c->kern_ns = 1e6 * (double)ms;
	c->kern_launches = 1;
	c->last_fallback = 0;
	if (ns)
		*ns = 1e6 * (double)ms;
	return GRAYHIP_OK;
}

static int make_rule(grayhip_ctx *c, double dt, size_t n_sub, size_t n_dump, StepRule *sr)
{
	if (!(dt != 0.0) || !std::isfinite(dt))
		return fail(c, GRAYHIP_EINVAL, "dt must be finite and non-zero");
	if (n_sub == 0 || n_sub > (size_t)INT_MAX)
		return fail(c, GRAYHIP_EINVAL, "n_sub must be in [1, INT_MAX]");
	if (n_dump == 0 || n_dump > (size_t)INT_MAX)
		return fail(c, GRAYHIP_EINVAL, "n_dump must be in [1, INT_MAX]");
	sr->dt = dt;
	sr->dts = dt / (double)n_sub;
	sr->n = (int)((INT_MAX / (long)n_sub) * (long)n_sub);
	sr->n_dump = (int)n_dump;
	return GRAYHIP_OK;
}

/* FP32 form: persistent pool passes over bands of image rows, nearest the
 * centre row first, each followed at once by the FP64 re-integration of its
 * hand-overs on a stream of its own, beside the later passes.  The bands
 * are |beta| <= 2 M, which holds the rays that wind closest to a prograde
 * photon orbit (at a = 0.998, i = 85 up to 1e6 steps, near beta = +-0.9,
 * alpha = -2.1), so their long FP64 chains start after a short pass; then
 * |beta| <= 8 M, the rest of the photon ring and the shadow edge for every
 * spin and inclination (critical impact parameters 2-7 M); then the outer
 * rows, which hand over few rays.  A band with fewer rays than the pool has
 * lanes would leave the GPU part idle for its whole pass: the first band is
 * widened to the rows that fill the lanes (|beta| <= 6 M at 1024^2), and a
 * later band that would still be short joins the next one.
 * Fallback launches cover the largest possible queue and idle threads
 * exit, so nothing waits on the host. */
static int run_fp32(grayhip_ctx *c, const StepRule &sr, double *ns)
{
	const int W = (int)c->cfg.w_rays, H = (int)c->cfg.h_rays;
	const double px = (double)H / fabs(c->p.h_img > 0.0 ? c->p.h_img : 1.0);
	const size_t lanes = (size_t)c->pool_blocks * 256;
	/* the first band is widened to the rows that fill the pool's lanes */
	const double fill = 0.5 * ceil((double)lanes / (double)W) / px;
	const double band[FP32_PASSES - 1] = {fill > 2.0 ? fill : 2.0, 8.0};
	/* row bands [lo, hi) of the passes before the outer one, innermost first */
	int lo[FP32_PASSES - 1], hi[FP32_PASSES - 1], n_in = 0;
	int plo = (H + 1) / 2, phi = (H + 1) / 2;   /* rows already taken: none */
	for (int k = 0; k < FP32_PASSES - 1; ++k) {
		int a = (int)floor(0.5 * H - band[k] * px), b = (int)ceil(0.5 * H + band[k] * px);
		a = a < 0 ? 0 : a;
		b = b > H ? H : b;
		const size_t rays = (size_t)((plo - a) + (b - phi)) * (size_t)W;
		if (a >= plo && b <= phi)
			continue;
		if (rays < lanes)
			continue;   /* too few rays for a pass of its own: join the next band */
		lo[n_in] = a;
		hi[n_in] = b;
		++n_in;
		plo = a;
		phi = b;
	}
	const int n_pass = n_in + 1;
	RayRange rr[FP32_PASSES];
	for (int k = 0; k < n_in; ++k) {
		const int ilo = k ? lo[k - 1] : (H + 1) / 2, ihi = k ? hi[k - 1] : (H + 1) / 2;
		rr[k] = {lo[k] * W, ilo * W, ihi * W, hi[k] * W};
	}
	rr[n_in] = {0, plo * W, phi * W, H * W};
	unsigned cnt[FP32_PASSES];
	Fallback *fb[FP32_PASSES];
	size_t off = 0;
	for (int k = 0; k < n_pass; ++k) {
		cnt[k] = (unsigned)((rr[k].e0 - rr[k].b0) + (rr[k].e1 - rr[k].b1));
		fb[k] = c->d_fb + off;
		off += cnt[k];
	}
	if (off != c->n_rays)
		return fail(c, GRAYHIP_EDEVICE, "pass ranges do not cover the image");
	/* Pass 0 runs on A; the later passes run in order on C beside it, so
	 * their blocks take the slots pass 0's finished blocks leave (pass 0
	 * has about one ray per lane and a tail of long rays).  Pass 0's
	 * hand-overs are re-integrated on B as soon as it ends, a middle pass's
	 * on A, the last pass's on C after it.
	 * events: 0/1 render, 2+2k / 3+2k pass k, 8 B done, 9 C done, 11 ready */
	hipStream_t A = c->stream, B = c->stream2, C = c->stream3;
	HCK(c, hipMemsetAsync(c->d_ctr, 0, FP32_PASSES * sizeof(Counters), A));
	HCK(c, hipEventRecord(c->ev[0], A));
	HCK(c, hipEventRecord(c->ev[11], A));
	HCK(c, hipEventRecord(c->ev[2], A));
	HCK(c, launch_pool_f32(c->p, sr, c->d_data, c->d_info, c->d_steps, c->n_rays, rr[0],
	                       c->d_ctr, fb[0], (unsigned)c->pool_blocks, A));
	HCK(c, hipEventRecord(c->ev[3], A));
	if (n_pass == 1) {
		HCK(c, launch_fallback_f64(c->p, sr, c->d_data, c->d_info, c->d_steps, c->n_rays, fb[0],
		                           c->d_ctr, cnt[0], A));
	} else {
		HCK(c, hipStreamWaitEvent(B, c->ev[3], 0));
		HCK(c, launch_fallback_f64(c->p, sr, c->d_data, c->d_info, c->d_steps, c->n_rays, fb[0],
		                           c->d_ctr, cnt[0], B));
		HCK(c, hipEventRecord(c->ev[8], B));
		HCK(c, hipStreamWaitEvent(C, c->ev[11], 0));
		for (int k = 1; k < n_pass; ++k) {
			HCK(c, hipEventRecord(c->ev[2 + 2 * k], C));
			HCK(c, launch_pool_f32(c->p, sr, c->d_data, c->d_info, c->d_steps, c->n_rays, rr[k],
			                       c->d_ctr + k, fb[k], (unsigned)c->pool_blocks, C));
			HCK(c, hipEventRecord(c->ev[3 + 2 * k], C));
			hipStream_t f = C;
			if (k < n_pass - 1) {
				f = A;
				HCK(c, hipStreamWaitEvent(A, c->ev[3 + 2 * k], 0));
			}
			HCK(c, launch_fallback_f64(c->p, sr, c->d_data, c->d_info, c->d_steps, c->n_rays,
			                           fb[k], c->d_ctr + k, cnt[k], f));
		}
		HCK(c, hipEventRecord(c->ev[9], C));
		HCK(c, hipStreamWaitEvent(A, c->ev[8], 0));
		HCK(c, hipStreamWaitEvent(A, c->ev[9], 0));
	}
	HCK(c, hipEventRecord(c->ev[1], A));
	Counters h[FP32_PASSES];
	HCK(c, hipMemcpyAsync(h, c->d_ctr, n_pass * sizeof(Counters), hipMemcpyDeviceToHost, A));
	HCK(c, hipStreamSynchronize(A));
	for (int k = 0; k < n_pass; ++k)
		if (h[k].n_fb > cnt[k])
			return fail(c, GRAYHIP_EDEVICE, "fallback queue overflow");
	float ms = 0.f;
	HCK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
	/* pool time: from the first pass's start to the last pass's end (the
	 * passes overlap) */
	double kern = 0.0;
	c->kern_launches = 0;
	c->last_fallback = 0;
	memset(&c->last_ctr, 0, sizeof(c->last_ctr));
	for (int k = 0; k < n_pass; ++k) {
		float kk = 0.f;
		HCK(c, hipEventElapsedTime(&kk, c->ev[2], c->ev[3 + 2 * k]));
		kern = (double)kk > kern ? (double)kk : kern;
		c->kern_launches += cnt[k] > 0;
		c->last_fallback += (int64_t)h[k].n_fb;
		c->last_ctr.n_fb += h[k].n_fb;
		c->last_ctr.pool_steps += h[k].pool_steps;
		c->last_ctr.fb_steps += h[k].fb_steps;
	}
	c->kern_ns = 1e6 * kern;
	if (ns)
		*ns = 1e6 * (double)ms;
	return GRAYHIP_OK;
}

/* Plain FP64 form: every interval in one launch (k_evolve_f64's flat loop;
 * the same bits as one launch per interval). */
static int run_fp64(grayhip_ctx *c, const StepRule &one, size_t n_dump, double *ns)
{
	StepRule sr = one;
	sr.n_dump = (int)n_dump;
	HCK(c, hipEventRecord(c->ev[0], c->stream));
	HCK(c, launch_evolve_f64(c->p, sr, c->d_data, c->d_info, c->d_steps, c->n_rays,
	                         c->stream));
	HCK(c, hipEventRecord(c->ev[1], c->stream));
	HCK(c, hipStreamSynchronize(c->stream));
	float ms = 0.f;
	HCK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
	c->kern_ns = 1e6 * (double)ms;
	c->kern_launches = 1;
	c->last_fallback = 0;
	if (ns)
		*ns = 1e6 * (double)ms;
	return GRAYHIP_OK;
}

extern "C" int grayhip_evolve(grayhip_ctx *c, double dt, size_t n_sub, double *ns)
{
	return grayhip_render(c, dt, n_sub, 1, ns);
}

extern "C" int grayhip_render(grayhip_ctx *c, double dt, size_t n_sub, size_t n_dump, double *ns)
{
	if (!c)
		return fail(nullptr, GRAYHIP_EINVAL, "null context");
	if (!c->have_icond)
		return fail(c, GRAYHIP_ESTATE, "grayhip_icond (or grayhip_write) must come first");
	StepRule sr;
	int rc = make_rule(c, dt, n_sub, n_dump, &sr);
	if (rc)
		return rc;
	HCK(c, hipSetDevice(c->cfg.device));
	if (c->st) {
		sr.n_dump = 1;
		return run_grid(c, sr, n_dump, ns);
	}
	if (c->cfg.precision == GRAYHIP_FP64) {
		sr.n_dump = 1;
		return run_fp64(c, sr, n_dump, ns);
	}
	return run_fp32(c, sr, ns);
}

extern "C" int grayhip_read(grayhip_ctx *c, void *data, void *info)
{
	if (!c)
		return fail(nullptr, GRAYHIP_EINVAL, "null context");
	const int sz = c->cfg.precision;
	HCK(c, hipSetDevice(c->cfg.device));
	if (data) {
		HCK(c, launch_soa_to_aos(c->d_data, c->d_tmp, sz, c->n_rays, c->nd, c->stream));
		HCK(c, hipMemcpyAsync(data, c->d_tmp, c->n_rays * (size_t)c->nd * sz,
		                      hipMemcpyDeviceToHost, c->stream));
		HCK(c, hipStreamSynchronize(c->stream));
	}
	if (info) {
		HCK(c, launch_cast(c->d_info, c->d_tmp, sz, c->n_rays, c->stream));
		HCK(c, hipMemcpyAsync(info, c->d_tmp, c->n_rays * (size_t)sz, hipMemcpyDeviceToHost,
		                      c->stream));
		HCK(c, hipStreamSynchronize(c->stream));
	}
	return GRAYHIP_OK;
}

extern "C" int grayhip_read_bl(grayhip_ctx *c, double r_match, double *out)
{
	if (!c || !out)
		return fail(c, GRAYHIP_EINVAL, "null argument");
	if (c->st)
		return fail(c, GRAYHIP_EINVAL, "the Boyer-Lindquist view needs the analytic spacetime");
	if (!(r_match > c->p.r_h))
		return fail(c, GRAYHIP_EINVAL, "r_match must lie outside the horizon");
	HCK(c, hipSetDevice(c->cfg.device));
	HCK(c, launch_to_bl(c->p.a, r_match, c->d_data, (double *)c->d_tmp, c->n_rays, c->stream));
	HCK(c, hipMemcpyAsync(out, c->d_tmp, c->n_rays * 8 * sizeof(double), hipMemcpyDeviceToHost,
	                      c->stream));
	HCK(c, hipStreamSynchronize(c->stream));
	return GRAYHIP_OK;
}

extern "C" int grayhip_read_covariant(grayhip_ctx *c, double *out)
{
	if (!c || !out)
		return fail(c, GRAYHIP_EINVAL, "null argument");
	HCK(c, hipSetDevice(c->cfg.device));
	if (c->st)
		HCK(c, launch_grid_covariant(grid_view(c), c->d_data, (double *)c->d_tmp, c->n_rays,
		                             c->stream));
	else
		HCK(c, launch_covariant(c->p.a, c->d_data, (double *)c->d_tmp, c->n_rays, c->stream));
	HCK(c, hipMemcpyAsync(out, c->d_tmp, c->n_rays * 8 * sizeof(double), hipMemcpyDeviceToHost,
	                      c->stream));
	HCK(c, hipStreamSynchronize(c->stream));
	return GRAYHIP_OK;
}

extern "C" int grayhip_read_f64(grayhip_ctx *c, double *data, double *info)
{
	if (!c)
		return fail(nullptr, GRAYHIP_EINVAL, "null context");
	HCK(c, hipSetDevice(c->cfg.device));
	if (data) {
		HCK(c, launch_soa_to_aos(c->d_data, c->d_tmp, 8, c->n_rays, c->nd, c->stream));
		HCK(c, hipMemcpyAsync(data, c->d_tmp, c->n_rays * (size_t)c->nd * 8,
		                      hipMemcpyDeviceToHost, c->stream));
	}
	if (info)
		HCK(c, hipMemcpyAsync(info, c->d_info, c->n_rays * 8, hipMemcpyDeviceToHost, c->stream));
	HCK(c, hipStreamSynchronize(c->stream));
	return GRAYHIP_OK;
}

/* The disk and uniform fluids absorb as a fixed power of frequency
 * (alpha_i shift = 0.1 nu_i^-5/2 n shift^-3/2), so every channel's depth is
 * tau_i = d_i T, d_i = 0.1 nu_i^-5/2, with one shared integral T; both forms
 * integrate T alone, taken from tau_0 (rt_disk.h, step_f32.h).  A restart
 * must therefore hold depths in that proportion: tau_i d_0 = tau_0 d_i to a
 * few rounding errors of the dump's real size. */
static int check_depths(grayhip_ctx *c, const void *data)
{
	const int nf = c->p.n_freq;
	if (c->st || nf < 2 ||
	    (c->p.fluid != GRAYHIP_FLUID_DISK && c->p.fluid != GRAYHIP_FLUID_UNIFORM))
		return GRAYHIP_OK;
	double d[GRAYHIP_MAX_FREQ];
	for (int i = 0; i < nf; ++i) {
		const double nu = c->p.nu[i];
		d[i] = 0.1 / (nu * nu * sqrt(nu));
	}
	const int sz = c->cfg.precision, nd = c->nd;
	const double tol = sz == 8 ? 64.0 * 2.220446049250313e-16 : 64.0 * 1.1920929e-7;
	for (size_t g = 0; g < c->n_rays; ++g) {
		double tau[GRAYHIP_MAX_FREQ];
		for (int i = 0; i < nf; ++i) {
			const size_t k = g * (size_t)nd + 8 + (size_t)nf + (size_t)i;
			tau[i] = sz == 8 ? ((const double *)data)[k] : (double)((const float *)data)[k];
		}
		for (int i = 1; i < nf; ++i) {
			const double lhs = tau[i] * d[0], rhs = tau[0] * d[i];
			if (fabs(lhs - rhs) > tol * fmax(fabs(lhs), fabs(rhs)))
				return fail(c, GRAYHIP_EINVAL,
				            "ray %zu: optical depth of channel %d (%.17g) is not "
				            "tau_0 (nu_0/nu_%d)^(5/2) = %.17g; with the disk and uniform "
				            "fluids every channel's depth is one shared integral",
				            g, i, tau[i], i, rhs / d[0]);
		}
	}
	return GRAYHIP_OK;
}

extern "C" int grayhip_write(grayhip_ctx *c, const void *data)
{
	if (!c || !data)
		return fail(c, GRAYHIP_EINVAL, "null argument");
	const int sz = c->cfg.precision;
	int rc = check_depths(c, data);
	if (rc)
		return rc;
	HCK(c, hipSetDevice(c->cfg.device));
	HCK(c, hipMemcpyAsync(c->d_tmp, data, c->n_rays * (size_t)c->nd * sz, hipMemcpyHostToDevice,
	                      c->stream));
	HCK(c, launch_aos_to_soa(c->d_tmp, c->d_data, sz, c->n_rays, c->nd, c->stream));
	if (c->st)
		HCK(c, launch_grid_info_f64(grid_view(c), c->d_data, c->d_info, c->n_rays, c->stream));
	else
		HCK(c, launch_info_f64(c->p, c->d_data, c->d_info, c->n_rays, c->stream));
	HCK(c, hipMemsetAsync(c->d_steps, 0, c->n_rays * sizeof(int64_t), c->stream));
	HCK(c, hipStreamSynchronize(c->stream));
	c->have_icond = true;
	return GRAYHIP_OK;
}

extern "C" int grayhip_read_steps(grayhip_ctx *c, int64_t *steps)
{
	if (!c || !steps)
		return fail(c, GRAYHIP_EINVAL, "null argument");
	HCK(c, hipSetDevice(c->cfg.device));
	HCK(c, hipMemcpy(steps, c->d_steps, c->n_rays * sizeof(int64_t), hipMemcpyDeviceToHost));
	return GRAYHIP_OK;
}

extern "C" int grayhip_stats(grayhip_ctx *c, int64_t *steps_total, int64_t *fallback_rays)
{
	if (!c)
		return fail(nullptr, GRAYHIP_EINVAL, "null context");
	if (steps_total) {
		std::vector<int64_t> s(c->n_rays);
		int rc = grayhip_read_steps(c, s.data());
		if (rc)
			return rc;
		int64_t t = 0;
		for (int64_t v : s)
			t += v;
		*steps_total = t;
	}
	if (fallback_rays)
		*fallback_rays = c->last_fallback;
	return GRAYHIP_OK;
}

extern "C" int grayhip_pass_steps(grayhip_ctx *c, int64_t *pool_steps, int64_t *fallback_steps)
{
	if (!c)
		return fail(nullptr, GRAYHIP_EINVAL, "null context");
	if (pool_steps)
		*pool_steps = (int64_t)c->last_ctr.pool_steps;
	if (fallback_steps)
		*fallback_steps = (int64_t)c->last_ctr.fb_steps;
	return GRAYHIP_OK;
}

extern "C" int grayhip_kernel_time(grayhip_ctx *c, double *ns, int *launches)
{
	if (!c)
		return fail(nullptr, GRAYHIP_EINVAL, "null context");
	if (ns)
		*ns = c->kern_ns;
	if (launches)
		*launches = c->kern_launches;
	return GRAYHIP_OK;
}

/* ---------------------------------------------------------------------
 * grayhip_render_static: icond + render for callers that cannot allocate.
 *
 * The FP32 form's scratch (pass counters and the hand-over queue of the
 * FP64 re-integration) lives in `scratch`, a file-scope device array of
 * fixed size, instead of memory a context allocates.  Each GPU holds its
 * own copy until the process ends; its device address is looked up once
 * per device (hipGetSymbolAddress with that device current) and only that
 * address is handed to kernels.  Every call on one GPU shares the array:
 * calls that overlap on one GPU race, and the counters are cleared by a
 * kernel at the start of every call.
 * --------------------------------------------------------------------- */

constexpr size_t SCRATCH_FB_OFF = 256;   /* Counters first, queue after */
constexpr size_t SCRATCH_BYTES =
    SCRATCH_FB_OFF + (size_t)GRAYHIP_STATIC_MAX_RAYS * sizeof(Fallback);
static_assert(sizeof(Counters) <= SCRATCH_FB_OFF, "counters overlap the queue");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int STATIC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[STATIC_MAX_DEVICES];

/* Device address of this device's copy of `scratch` (device current). */
static int static_scratch(int device, unsigned char **out)
{
	if (device < 0 || device >= STATIC_MAX_DEVICES)
		return fail(nullptr, GRAYHIP_EINVAL, "device %d beyond the %d the library serves",
		            device, STATIC_MAX_DEVICES);
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_ptr[device]) {
		void *ptr = nullptr;
		hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr)
			return fail(nullptr, GRAYHIP_EDEVICE, "hipGetSymbolAddress(scratch): %s",
			            hipGetErrorString(e));
		g_scratch_ptr[device] = ptr;
	}
	*out = (unsigned char *)g_scratch_ptr[device];
	return GRAYHIP_OK;
}

#define SCKG(expr)                                                                        \
	do {                                                                              \
		hipError_t e_ = (expr);                                                   \
		if (e_ != hipSuccess)                                                     \
			return fail(nullptr, GRAYHIP_EDEVICE, "%s: %s", #expr,            \
			            hipGetErrorString(e_));                               \
	} while (0)

extern "C" int grayhip_render_static(const grayhip_config *cfg, double w_img, double h_img,
                                     double r_obs, double i_obs, double j_obs, double dt,
                                     size_t n_sub, size_t n_dump, double *data, double *info,
                                     int64_t *steps, void *stream)
{
	if (!cfg || !data || !info || !steps)
		return fail(nullptr, GRAYHIP_EINVAL, "null argument");
	int rc = check_config(cfg);
	if (rc)
		return rc;
	if (cfg->spacetime)
		return fail(nullptr, GRAYHIP_EINVAL, "grayhip_render_static renders analytic Kerr only");
	const size_t n = cfg->w_rays * cfg->h_rays;
	if (n > (size_t)GRAYHIP_STATIC_MAX_RAYS)
		return fail(nullptr, GRAYHIP_EINVAL, "%zu rays exceed the static scratch's %d",
		            n, GRAYHIP_STATIC_MAX_RAYS);
	if (!(r_obs > 0.0))
		return fail(nullptr, GRAYHIP_EINVAL, "r_obs must be positive");
	StepRule sr;
	if ((rc = make_rule(nullptr, dt, n_sub, n_dump, &sr)))
		return rc;
	RayParams p;
	fill_params(cfg, p);
	fill_camera(p, w_img, h_img, r_obs, i_obs, j_obs);
	hipStream_t st = (hipStream_t)stream;
	SCKG(hipSetDevice(cfg->device));
	unsigned char *base = nullptr;
	if ((rc = static_scratch(cfg->device, &base)))
		return rc;
	Counters *ctr = (Counters *)base;
	Fallback *fb = (Fallback *)(base + SCRATCH_FB_OFF);

	SCKG(launch_icond(p, data, info, steps, n, st));
	if (cfg->precision == GRAYHIP_FP64) {
		sr.n_dump = (int)n_dump;
		SCKG(launch_evolve_f64(p, sr, data, info, steps, n, st));
		return GRAYHIP_OK;
	}
	/* one pool pass over every ray, then the FP64 re-integration of its
	 * hand-overs (at most one per ray: the queue holds n entries) */
	int bpc = 0, n_cu = 0;
	SCKG(pool_f32_blocks_per_cu(p.n_freq, &bpc));
	SCKG(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, cfg->device));
	const size_t need = (n + 255) / 256, full = (size_t)(bpc > 0 ? bpc : 1) * (size_t)n_cu;
	const unsigned blocks = (unsigned)(need < full ? need : full);
	const RayRange rr = {0, (int)n, (int)n, (int)n};
	SCKG(launch_zero_counters(ctr, 1, st));
	SCKG(launch_pool_f32(p, sr, data, info, steps, n, rr, ctr, fb, blocks, st));
	SCKG(launch_fallback_f64(p, sr, data, info, steps, n, fb, ctr, (unsigned)n, st));
	return GRAYHIP_OK;
}
