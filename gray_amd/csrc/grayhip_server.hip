/*
 * grayhip_server.hip -- the serving pool's C ABI (include/grayhip.h,
 * grayhip_server_*) and its HIP back-end.
 *
 * The host scheduler (srv_sched.h) decides admissions and each epoch's
 * queues; the back-end here owns the device memory (request table, Kerr
 * classes, queue segments, the two resume, FP64 and tile lists that swap
 * roles every epoch, the output arena) and runs an epoch as three
 * persistent kernels on three streams side by side: FP32 Kerr pool, FP64
 * re-integration, gridded tiles (kernels_serve.hip).
 */
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "launch_serve.h"
#include "step_f32.h"
#include "spacetime.h"
#include "srv_sched.h"

using namespace gray;

namespace {

constexpr int SRV_FB_CAP = 1 << 19;     /* entries of each FP64 list          */
constexpr int SRV_MAX_SEGS = 1 << 15;   /* fresh segments per kind per epoch  */
/* target epoch duration: the longest a request that arrives waits to
 * start; 8 against 50 ms measured 3.65e10 against 3.67e10 ray-steps/s on
 * the 32-request stream sample */
constexpr double SRV_EPOCH_MS = 8.0;

/* Prograde ISCO (Bardeen, Press & Teukolsky 1972), M = 1. */
double r_isco_host(double a)
{
	const double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
	const double z2 = sqrt(3.0 * a * a + z1 * z1);
	return 3.0 + z2 - sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

class HipDevice : public srv::Device {
public:
	grayhip_server_config cfg;
	srv::Caps caps;
	int clock_khz = 0;    /* wall clock ticks per ms (wall_clock64)          */
	hipStream_t s[3] = {nullptr, nullptr, nullptr};
	hipEvent_t ev[7] = {};
	SrvRequest *d_reqs = nullptr;
	unsigned *d_remaining = nullptr;
	SrvClass *d_cls = nullptr;
	int *d_kseg = nullptr, *d_gseg = nullptr;        /* tab, first, start (3 x max) */
	SrvResume res[2] = {};
	SrvFb fb[2] = {};
	SrvTiles tiles[2] = {};
	SrvCtr *d_ctr = nullptr;
	int *d_done = nullptr;
	char *arena = nullptr;
	/* pinned staging */
	SrvCtr *h_ctr = nullptr;
	int *h_done = nullptr;
	volatile unsigned *h_started = nullptr;   /* written by the GPU (coherent) */
	unsigned epoch_no = 0;
	SrvClass *h_cls = nullptr;
	int *h_kseg = nullptr, *h_gseg = nullptr;
	/* messages per calling thread: an epoch and a result copy may fail on
	 * different threads at once */
	static char *err_buf()
	{
		static thread_local char err[256];
		return err;
	}

	~HipDevice() override { release(); }

	int fail(int code, const char *fmt, ...)
	{
		va_list ap;
		va_start(ap, fmt);
		vsnprintf(err_buf(), 256, fmt, ap);
		va_end(ap);
		return code;
	}
	const char *error() const override { return err_buf(); }

	int init(const grayhip_server_config &c)
	{
		cfg = c;
		hipError_t e = hipSetDevice(c.device);
		int n_cu = 0, kb = 0, gb = 0, kbh = 0;
		if (e == hipSuccess)
			e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, c.device);
		if (e == hipSuccess)
			e = srv_kerr_blocks_per_cu(&kb);
		if (e == hipSuccess)
			e = srv_grid_blocks_per_cu(&gb);
		if (e == hipSuccess)
			e = srv_kerr_beside_heavy(&kbh);
		if (e == hipSuccess)
			e = hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, c.device);
		if (e == hipSuccess && clock_khz <= 0)
			return fail(GRAYHIP_EDEVICE, "device reports no wall clock rate");
		if (e != hipSuccess)
			return fail(GRAYHIP_EDEVICE, "device query: %s", hipGetErrorString(e));
		/* `slots` bounds the rays in flight of each launch */
		const int want = (int)std::min<size_t>((c.slots + SRV_BLOCK - 1) / SRV_BLOCK, 1u << 20);
		caps.kerr_blocks = std::max(1, std::min(std::max(kb, 1) * n_cu, want));
		caps.grid_blocks = std::max(1, std::min(std::max(gb, 1) * n_cu, want));
		caps.n_cu = std::max(1, n_cu);
		caps.kerr_beside = std::max(0, std::min(kbh, std::max(kb, 1)));
		caps.fb_cap = SRV_FB_CAP;
		caps.max_segs = SRV_MAX_SEGS;
		caps.max_requests = c.max_requests;
		caps.arena = c.hbm_budget;
		const size_t R = (size_t)c.max_requests;
		/* an epoch writes at most max(lanes, entries resumed) per class (or
		 * waves, records for the grid): twice a launch's lanes bound it */
		const int kcap = 2 * (caps.kerr_blocks + SRV_MAX_CLASSES) * SRV_BLOCK;
		const int gcap = 2 * (caps.grid_blocks * SRV_GWAVES + SRV_GWAVES);
		caps.res_cap = kcap;
		caps.tile_cap = gcap;
		struct { void **p; size_t b; } al[] = {
			{(void **)&d_reqs, R * sizeof(SrvRequest)},
			{(void **)&d_remaining, R * sizeof(unsigned)},
			{(void **)&d_cls, SRV_MAX_CLASSES * sizeof(SrvClass)},
			{(void **)&d_kseg, 3 * (size_t)SRV_MAX_SEGS * sizeof(int)},
			{(void **)&d_gseg, 3 * (size_t)SRV_MAX_SEGS * sizeof(int)},
			{(void **)&res[0].meta, (size_t)kcap * sizeof(int4)},
			{(void **)&res[1].meta, (size_t)kcap * sizeof(int4)},
			{(void **)&res[0].steps, (size_t)kcap * sizeof(int64_t)},
			{(void **)&res[1].steps, (size_t)kcap * sizeof(int64_t)},
			{(void **)&res[0].st, (size_t)kcap * 2 * SRV_NK * sizeof(float)},
			{(void **)&res[1].st, (size_t)kcap * 2 * SRV_NK * sizeof(float)},
			{(void **)&fb[0].meta, (size_t)SRV_FB_CAP * sizeof(int4)},
			{(void **)&fb[1].meta, (size_t)SRV_FB_CAP * sizeof(int4)},
			{(void **)&fb[0].steps, (size_t)SRV_FB_CAP * sizeof(int64_t)},
			{(void **)&fb[1].steps, (size_t)SRV_FB_CAP * sizeof(int64_t)},
			{(void **)&fb[0].x, (size_t)SRV_FB_CAP * SRV_NC * sizeof(double)},
			{(void **)&fb[1].x, (size_t)SRV_FB_CAP * SRV_NC * sizeof(double)},
			{(void **)&tiles[0].hdr, (size_t)gcap * sizeof(int2)},
			{(void **)&tiles[1].hdr, (size_t)gcap * sizeof(int2)},
			{(void **)&tiles[0].brick, (size_t)gcap * sizeof(int4)},
			{(void **)&tiles[1].brick, (size_t)gcap * sizeof(int4)},
			{(void **)&tiles[0].meta, (size_t)gcap * 64 * sizeof(int4)},
			{(void **)&tiles[1].meta, (size_t)gcap * 64 * sizeof(int4)},
			{(void **)&tiles[0].steps, (size_t)gcap * 64 * sizeof(int64_t)},
			{(void **)&tiles[1].steps, (size_t)gcap * 64 * sizeof(int64_t)},
			{(void **)&tiles[0].st, (size_t)gcap * 64 * 2 * SRV_NG * sizeof(float)},
			{(void **)&tiles[1].st, (size_t)gcap * 64 * 2 * SRV_NG * sizeof(float)},
			{(void **)&d_ctr, sizeof(SrvCtr)},
			{(void **)&d_done, R * sizeof(int)},
			{(void **)&arena, c.hbm_budget},
		};
		for (auto &a : al)
			if (e == hipSuccess)
				e = hipMalloc(a.p, a.b);
		if (e == hipSuccess)
			e = hipHostMalloc((void **)&h_ctr, sizeof(SrvCtr), 0);
		if (e == hipSuccess)
			e = hipHostMalloc((void **)&h_started, sizeof(unsigned), hipHostMallocCoherent);
		if (e == hipSuccess)
			e = hipHostMalloc((void **)&h_done, R * sizeof(int), 0);
		if (e == hipSuccess)
			e = hipHostMalloc((void **)&h_cls, SRV_MAX_CLASSES * sizeof(SrvClass), 0);
		if (e == hipSuccess)
			e = hipHostMalloc((void **)&h_kseg, 3 * (size_t)SRV_MAX_SEGS * sizeof(int), 0);
		if (e == hipSuccess)
			e = hipHostMalloc((void **)&h_gseg, 3 * (size_t)SRV_MAX_SEGS * sizeof(int), 0);
		for (int i = 0; i < 3 && e == hipSuccess; ++i)
			e = hipStreamCreateWithFlags(&s[i], hipStreamNonBlocking);
		for (int i = 0; i < 7 && e == hipSuccess; ++i)
			e = hipEventCreate(&ev[i]);
		if (e != hipSuccess)
			return fail(GRAYHIP_ENOMEM, "server allocation: %s", hipGetErrorString(e));
		for (int k = 0; k < 2; ++k) {
			res[k].cap = kcap;
			fb[k].cap = SRV_FB_CAP;
			tiles[k].cap = gcap;
		}
		return GRAYHIP_OK;
	}

	void release()
	{
		(void)hipSetDevice(cfg.device);
		for (int i = 0; i < 3; ++i)
			if (s[i])
				(void)hipStreamDestroy(s[i]);
		for (int i = 0; i < 7; ++i)
			if (ev[i])
				(void)hipEventDestroy(ev[i]);
		void *dev[] = {d_reqs, d_remaining, d_cls, d_kseg, d_gseg, d_ctr, d_done, arena};
		for (void *p : dev)
			(void)hipFree(p);
		for (int k = 0; k < 2; ++k) {
			(void)hipFree(res[k].meta);
			(void)hipFree(res[k].steps);
			(void)hipFree(res[k].st);
			(void)hipFree(fb[k].meta);
			(void)hipFree(fb[k].steps);
			(void)hipFree(fb[k].x);
			(void)hipFree(tiles[k].hdr);
			(void)hipFree(tiles[k].brick);
			(void)hipFree(tiles[k].meta);
			(void)hipFree(tiles[k].steps);
			(void)hipFree(tiles[k].st);
		}
		void *host[] = {h_ctr, h_done, h_cls, h_kseg, h_gseg, (void *)h_started};
		for (void *p : host)
			(void)hipHostFree(p);
	}

	int admit(int tab, const srv::Req &r) override
	{
		const grayhip_request *q = &r.q;
		SrvRequest D;
		memset(&D, 0, sizeof(D));
		RayParams &p = D.p;
		const int nf = q->n_freq;
		p.a = q->spacetime ? 0.0 : q->a_spin;
		p.r_h = 1.0 + sqrt(1.0 - p.a * p.a);
		p.escape = q->escape;
		p.eps_stop = q->eps_stop > 0.0 ? q->eps_stop
		             : (q->escape == GRAYHIP_ESC_OUTWARD ? 0.01 * p.r_h : 0.01);
		p.r_esc = q->r_esc;
		p.r_isco = r_isco_host(p.a);
		p.n_freq = nf;
		p.fluid = nf > 0 ? q->fluid : GRAYHIP_FLUID_NONE;
		for (int i = 0; i < GRAYHIP_MAX_FREQ; ++i)
			p.nu[i] = nf > 0 ? q->nu[i < nf && i < SRV_NF ? i : 0] : 0.0;
		const double d2r = M_PI / 180.0;
		p.w_img = q->w_img;
		p.h_img = q->h_img;
		p.r_obs = q->r_obs;
		p.si = sin(d2r * q->i_obs);
		p.ci = cos(d2r * q->i_obs);
		p.sj = sin(d2r * q->j_obs);
		p.cj = cos(d2r * q->j_obs);
		p.w_rays = (int)q->w_rays;
		p.h_rays = (int)q->h_rays;
		StepRule &sr = D.sr;
		sr.dt = q->dt;
		sr.dts = q->dt / (double)q->n_sub;
		sr.n = (int)((INT32_MAX / (long)q->n_sub) * (long)q->n_sub);
		sr.n_dump = (int)q->n_dump;
		if (q->spacetime) {
			D.gv = spacetime_view(q->spacetime, q->t_init, 0, q->escape == GRAYHIP_ESC_REF)
			           .view<float>();
			rt_thermal_init(&D.rt, 4.3e6, 0.0, nf, q->nu);
		}
		D.kind = r.kind;
		D.nd = r.nd;
		D.cls = r.cls;
		D.tw = (int)((q->w_rays + 7) / 8);
		double *base = (double *)(arena + r.off);
		D.out_data = base;
		D.out_info = base + r.n_rays * (size_t)r.nd;
		D.out_steps = (int64_t *)(D.out_info + r.n_rays);
		const unsigned rem = (unsigned)r.n_rays;
		hipError_t e = hipSetDevice(cfg.device);
		if (e == hipSuccess)
			e = hipMemcpy(d_reqs + tab, &D, sizeof(D), hipMemcpyHostToDevice);
		if (e == hipSuccess)
			e = hipMemcpy(d_remaining + tab, &rem, sizeof(rem), hipMemcpyHostToDevice);
		if (e != hipSuccess)
			return fail(GRAYHIP_EDEVICE, "admit: %s", hipGetErrorString(e));
		return GRAYHIP_OK;
	}

	/* Segments of one kind into the pinned staging and the device: tab,
	 * first, and each segment's first queue item within its class. */
	static void pack_segs(const std::vector<srv::Seg> &segs, const srv::Plan &p, bool kerr, int *h)
	{
		const size_t n = segs.size();
		int *tab = h, *first = h + SRV_MAX_SEGS, *start = h + 2 * SRV_MAX_SEGS;
		for (size_t j = 0; j < n; ++j) {
			tab[j] = segs[j].tab;
			first[j] = segs[j].first;
		}
		if (kerr) {
			for (int k = 0; k < p.n_cls; ++k) {
				unsigned acc = 0;
				for (int j = p.cls[k].seg0; j < p.cls[k].seg1; ++j) {
					start[j] = (int)acc;
					acc += segs[(size_t)j].count;
				}
			}
		} else {
			unsigned acc = 0;
			for (size_t j = 0; j < n; ++j) {
				start[j] = (int)acc;
				acc += segs[j].count;
			}
		}
	}

	hipError_t put_segs(const int *h, int *d, size_t n, hipStream_t st)
	{
		hipError_t e = hipSuccess;
		for (int k = 0; k < 3 && e == hipSuccess && n > 0; ++k)
			e = hipMemcpyAsync(d + (size_t)k * SRV_MAX_SEGS, h + (size_t)k * SRV_MAX_SEGS,
			                   n * sizeof(int), hipMemcpyHostToDevice, st);
		return e;
	}

#define DCK(expr)                                                                         \
	do {                                                                              \
		hipError_t e_ = (expr);                                                   \
		if (e_ != hipSuccess)                                                     \
			return fail(GRAYHIP_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
	} while (0)

	int epoch(const srv::Plan &p, srv::Result &out) override
	{
		DCK(hipSetDevice(cfg.device));
		if (p.kseg.size() > (size_t)SRV_MAX_SEGS || p.gseg.size() > (size_t)SRV_MAX_SEGS)
			return fail(GRAYHIP_EDEVICE, "too many queue segments");
		/* the lists the kernels index */
		bool ok = p.kerr_blocks <= caps.kerr_blocks + SRV_MAX_CLASSES &&
		          p.grid_blocks <= caps.grid_blocks && p.fb_blocks <= caps.kerr_blocks &&
		          p.n_cls <= SRV_MAX_CLASSES && p.n_fb >= 0 && p.fb_off >= 0 &&
		          p.fb_off + p.n_fb <= fb[0].cap && p.g_resume >= 0 && p.g_res_off >= 0 &&
		          p.g_res_off + p.g_resume <= tiles[0].cap && p.g_out_cap <= tiles[0].cap &&
		          (p.grid_blocks == 0 || p.g_out_cap >= p.grid_blocks * SRV_GWAVES);
		for (int k = 0; k < p.n_cls && ok; ++k) {
			const srv::ClassPlan &cp = p.cls[k];
			ok = cp.block0 >= 0 && cp.block0 < cp.block1 && cp.block1 <= p.kerr_blocks &&
			     cp.res_off >= 0 && cp.res_off + cp.n_resume <= res[0].cap && cp.out_base >= 0 &&
			     cp.out_base + cp.out_cap <= res[0].cap &&
			     cp.out_cap >= (cp.block1 - cp.block0) * SRV_BLOCK && cp.out_cap >= cp.n_resume;
		}
		if (!ok)
			return fail(GRAYHIP_EDEVICE, "epoch plan beyond the pool's lists");
		const int in = p.flip, o = p.flip ^ 1;
		for (int k = 0; k < p.n_cls; ++k) {
			const srv::ClassPlan &cp = p.cls[k];
			SrvClass &C = h_cls[k];
			memset(&C, 0, sizeof(C));
			C.sr.dt = cp.key.dt;
			C.sr.dts = cp.key.dt / (double)cp.key.n_sub;
			C.sr.n = (int)((INT32_MAX / (long)cp.key.n_sub) * (long)cp.key.n_sub);
			C.sr.n_dump = (int)cp.key.n_dump;
			for (int i = 0; i < SRV_NF; ++i)
				C.nu[i] = cp.key.nu[i];
			/* the FP32 step's class constants, converted here so that the
			 * kernel holds them in scalar registers */
			C.n_over_dt = (double)C.sr.n / C.sr.dt;
			C.dt_over_n = C.sr.dt / (double)C.sr.n;
			C.dts = (float)C.sr.dts;
			if (cp.key.n_freq > 0) {
				const RtConst<SRV_NF> rc = make_rtconst<SRV_NF>(C);
				for (int i = 0; i < SRV_NF; ++i) {
					C.l2c[i] = rc.l2c[i];
					C.d2[i] = rc.d2[i];
				}
			}
			C.r_esc = cp.key.r_esc;
			C.n_freq = cp.key.n_freq;
			C.fluid = cp.key.n_freq > 0 ? cp.key.fluid : GRAYHIP_FLUID_NONE;
			C.escape = cp.key.escape;
			C.block0 = cp.block0;
			C.block1 = cp.block1;
			C.res_off = cp.res_off;
			C.n_resume = cp.n_resume;
			C.out_base = cp.out_base;
			C.out_cap = cp.out_cap;
			C.seg0 = cp.seg0;
			C.seg1 = cp.seg1;
			C.n_fresh = cp.n_fresh;
		}
		pack_segs(p.kseg, p, true, h_kseg);
		pack_segs(p.gseg, p, false, h_gseg);
		hipStream_t A = s[0];
		if (p.n_cls > 0)
			DCK(hipMemcpyAsync(d_cls, h_cls, (size_t)p.n_cls * sizeof(SrvClass),
			                   hipMemcpyHostToDevice, A));
		DCK(put_segs(h_kseg, d_kseg, p.kseg.size(), A));
		DCK(put_segs(h_gseg, d_gseg, p.gseg.size(), A));
		const unsigned long long budget =
		    p.epoch_ms > 0.0 ? (unsigned long long)(p.epoch_ms * clock_khz) : 0ull;
		DCK(launch_srv_clear(d_ctr, budget, A));
		DCK(hipEventRecord(ev[0], A));
		DCK(hipStreamWaitEvent(s[1], ev[0], 0));
		DCK(hipStreamWaitEvent(s[2], ev[0], 0));

		const bool run[3] = {p.n_cls > 0 && p.kerr_blocks > 0, p.n_fb > 0 && p.fb_blocks > 0,
		                     p.grid_blocks > 0 && (p.g_resume > 0 || p.g_fresh > 0)};
		SrvHeavy H;
		H.total = (run[1] ? (unsigned)p.fb_blocks : 0u) + (run[2] ? (unsigned)p.grid_blocks : 0u);
		H.epoch = ++epoch_no;
		H.started = h_started;

		SrvKerrArgs K;
		K.reqs = d_reqs;
		K.cls = d_cls;
		K.n_cls = p.n_cls;
		K.seg = SrvSegs{d_kseg, d_kseg + SRV_MAX_SEGS, (const unsigned *)(d_kseg + 2 * SRV_MAX_SEGS)};
		K.in = res[in];
		K.out = res[o];
		K.fb = fb[o];
		K.ctr = d_ctr;
		K.remaining = d_remaining;
		K.done = d_done;
		K.K = p.K_kerr;
		SrvFbArgs F;
		F.reqs = d_reqs;
		F.in = fb[in];
		F.in_off = p.fb_off;
		F.n_in = p.n_fb;
		F.out = fb[o];
		F.ctr = d_ctr;
		F.remaining = d_remaining;
		F.done = d_done;
		F.K = p.K_fb;
		F.heavy = H;
		SrvGridArgs Gd;
		Gd.reqs = d_reqs;
		Gd.seg = SrvSegs{d_gseg, d_gseg + SRV_MAX_SEGS, (const unsigned *)(d_gseg + 2 * SRV_MAX_SEGS)};
		Gd.seg0 = 0;
		Gd.seg1 = (int)p.gseg.size();
		Gd.res_off = p.g_res_off;
		Gd.n_resume = p.g_resume;
		Gd.n_fresh = p.g_fresh;
		Gd.out_cap = p.g_out_cap;
		Gd.in = tiles[in];
		Gd.out = tiles[o];
		Gd.ctr = d_ctr;
		Gd.remaining = d_remaining;
		Gd.done = d_done;
		Gd.K = p.K_grid;
		Gd.heavy = H;

		/* the FP64 and grid launches first; the Kerr launch once all their
		 * blocks are resident (or after a bound, past which the epoch's
		 * stop time covers a launch that could not fit) */
		for (int k = 1; k < 3; ++k) {
			if (!run[k])
				continue;
			DCK(hipEventRecord(ev[1 + 2 * k], s[k]));
			if (k == 1)
				DCK(launch_srv_fb64(F, p.fb_blocks, s[1]));
			else
				DCK(launch_srv_grid(Gd, p.grid_blocks, s[2]));
			DCK(hipEventRecord(ev[2 + 2 * k], s[k]));
		}
		if (run[0]) {
			if (H.total > 0) {
				const auto t0 = std::chrono::steady_clock::now();
				while (*h_started != H.epoch &&
				       std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(2))
					;
			}
			DCK(hipEventRecord(ev[1], s[0]));
			DCK(launch_srv_kerr(K, p.kerr_blocks, s[0]));
			DCK(hipEventRecord(ev[2], s[0]));
		}
		for (int k = 0; k < 3; ++k)
			DCK(hipStreamSynchronize(s[k]));
		DCK(hipMemcpy(h_ctr, d_ctr, sizeof(SrvCtr), hipMemcpyDeviceToHost));
		if (h_ctr->n_done > (unsigned)cfg.max_requests)
			return fail(GRAYHIP_EDEVICE, "done list overflowed");
		if (h_ctr->n_done)
			DCK(hipMemcpy(h_done, d_done, h_ctr->n_done * sizeof(int), hipMemcpyDeviceToHost));
		for (int k = 0; k < SRV_MAX_CLASSES; ++k) {
			out.res_claimed[k] = h_ctr->res_claimed[k];
			out.fresh_claimed[k] = h_ctr->fresh_claimed[k];
			out.saved[k] = h_ctr->saved[k];
		}
		out.fb_saved = h_ctr->fb_saved;
		out.g_res = h_ctr->g_res;
		out.g_fresh = h_ctr->g_fresh;
		out.g_saved = h_ctr->g_saved;
		out.overflow = h_ctr->overflow;
		out.done.assign(h_done, h_done + h_ctr->n_done);
		for (int k = 0; k < 3; ++k) {
			out.steps[k] = h_ctr->steps[k];
			out.ms[k] = -1.0;
			out.late[k] = h_ctr->late[k];
		}
		/* kernel times in scheduler order: Kerr, FP64, grid */
		for (int k = 0; k < 3; ++k) {
			float ms = 0.f;
			if (run[k] && hipEventElapsedTime(&ms, ev[1 + 2 * k], ev[2 + 2 * k]) == hipSuccess)
				out.ms[k] = (double)ms;
		}
		return GRAYHIP_OK;
	}

	int fetch(const srv::Req &r, double *data, double *info, int64_t *steps) override
	{
		DCK(hipSetDevice(cfg.device));
		const double *base = (const double *)(arena + r.off);
		if (data)
			DCK(hipMemcpy(data, base, r.n_rays * (size_t)r.nd * 8, hipMemcpyDeviceToHost));
		if (info)
			DCK(hipMemcpy(info, base + r.n_rays * (size_t)r.nd, r.n_rays * 8,
			              hipMemcpyDeviceToHost));
		if (steps)
			DCK(hipMemcpy(steps, base + r.n_rays * (size_t)(r.nd + 1), r.n_rays * 8,
			              hipMemcpyDeviceToHost));
		return GRAYHIP_OK;
	}
#undef DCK
};

} // namespace

struct grayhip_server {
	HipDevice dev;
	srv::Sched *sched = nullptr;
};

/* The message of the calling thread's last failed server call. */
static thread_local char t_err[256];

static int vfail(grayhip_server *, int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(t_err, sizeof(t_err), fmt, ap);
	va_end(ap);
	return code;
}

extern "C" const char *grayhip_server_error(const grayhip_server *)
{
	return t_err;
}

extern "C" void grayhip_server_destroy(grayhip_server *v)
{
	if (!v)
		return;
	delete v->sched;
	delete v;
}

extern "C" int grayhip_server_create(const grayhip_server_config *cfg, grayhip_server **out)
{
	if (!cfg || !out)
		return vfail(nullptr, GRAYHIP_EINVAL, "null argument");
	*out = nullptr;
	if (cfg->slots == 0 || cfg->slots > (1u << 28) || cfg->epoch_steps <= 0 ||
	    cfg->max_requests <= 0 || cfg->max_requests > (1 << 20) || cfg->hbm_budget == 0)
		return vfail(nullptr, GRAYHIP_EINVAL, "bad server configuration");
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
		return vfail(nullptr, GRAYHIP_ENODEV, "no HIP device visible");
	if (cfg->device < 0 || cfg->device >= ndev)
		return vfail(nullptr, GRAYHIP_EINVAL, "device %d out of range", cfg->device);
	grayhip_server *v = new (std::nothrow) grayhip_server();
	if (!v)
		return vfail(nullptr, GRAYHIP_ENOMEM, "host allocation failed");
	int rc = v->dev.init(*cfg);
	if (rc) {
		vfail(nullptr, rc, "%s", v->dev.error());
		delete v;
		return rc;
	}
	srv::Config sc;
	sc.epoch_steps = cfg->epoch_steps;
	sc.epoch_ms = SRV_EPOCH_MS;
	v->sched = new (std::nothrow) srv::Sched(sc, v->dev.caps, &v->dev);
	if (!v->sched) {
		delete v;
		return vfail(nullptr, GRAYHIP_ENOMEM, "host allocation failed");
	}
	*out = v;
	return GRAYHIP_OK;
}

extern "C" int grayhip_server_submit(grayhip_server *v, const grayhip_request *q, int64_t *id)
{
	if (!v || !q || !id)
		return vfail(v, GRAYHIP_EINVAL, "null argument");
	/* each side bounded first, so the product cannot wrap */
	if (q->w_rays == 0 || q->h_rays == 0 || q->w_rays > (size_t)INT32_MAX ||
	    q->h_rays > (size_t)INT32_MAX || q->w_rays * q->h_rays > (size_t)INT32_MAX)
		return vfail(v, GRAYHIP_EINVAL, "bad resolution");
	if (q->n_freq < 0 || q->n_freq > SRV_NF)
		return vfail(v, GRAYHIP_EINVAL, "n_freq must be in [0, %d]", SRV_NF);
	if (!(fabs(q->a_spin) < 1.0) || q->n_sub == 0 || q->n_sub > (size_t)INT32_MAX ||
	    q->n_dump == 0 || q->n_dump > (size_t)INT32_MAX || !(q->dt != 0.0) ||
	    !std::isfinite(q->dt))
		return vfail(v, GRAYHIP_EINVAL, "bad spin or schedule");
	if (q->escape != GRAYHIP_ESC_REF && q->escape != GRAYHIP_ESC_OUTWARD)
		return vfail(v, GRAYHIP_EINVAL, "unknown escape rule %d", q->escape);
	if (!(q->r_obs > 0.0))
		return vfail(v, GRAYHIP_EINVAL, "r_obs must be positive");
	for (int i = 0; i < q->n_freq; ++i)
		if (!(q->nu[i] > 0.0))
			return vfail(v, GRAYHIP_EINVAL, "frequency channel %d must be positive", i);
	if (q->spacetime) {
		if (q->spacetime->desc.dtype != 4 || q->spacetime->desc.device != v->dev.cfg.device)
			return vfail(v, GRAYHIP_EINVAL, "gridded requests need an FP32 spacetime on this "
			                                "device");
		if (q->n_freq > 0 && q->fluid != GRAYHIP_FLUID_THERMAL)
			return vfail(v, GRAYHIP_EINVAL, "transfer on a numerical spacetime uses the "
			                                "thermal fluid (the reference flow)");
		if (q->n_freq > 0 && !grayhip_spacetime_has_rho(q->spacetime))
			return vfail(v, GRAYHIP_ESTATE, "transfer needs the density of every level in "
			                                "use (grayhip_spacetime_upload_rho)");
	} else if (q->n_freq > 0 && q->fluid != GRAYHIP_FLUID_DISK &&
	           q->fluid != GRAYHIP_FLUID_UNIFORM) {
		return vfail(v, GRAYHIP_EINVAL, "served analytic requests use the disk or uniform fluid");
	}
	srv::Req r;
	memset(&r, 0, sizeof(r));
	r.q = *q;
	r.kind = q->spacetime ? SRV_GRID : SRV_KERR;
	r.nd = 8 + 2 * q->n_freq;
	r.n_rays = q->w_rays * q->h_rays;
	r.units = r.kind == SRV_KERR ? r.n_rays : ((q->w_rays + 7) / 8) * ((q->h_rays + 7) / 8);
	r.bytes = (r.n_rays * (size_t)(r.nd + 2) * 8 + 255) & ~(size_t)255;
	if (r.kind == SRV_KERR) {
		srv::ClassKey &k = r.key;
		k.dt = q->dt;
		k.n_sub = q->n_sub;
		k.n_dump = q->n_dump;
		k.n_freq = q->n_freq;
		k.fluid = q->n_freq > 0 ? q->fluid : GRAYHIP_FLUID_NONE;
		k.escape = q->escape;
		k.r_esc = q->r_esc;
		for (int i = 0; i < SRV_NF; ++i)
			k.nu[i] = q->n_freq > 0 ? q->nu[i < q->n_freq ? i : 0] : 0.0;
	}
	return v->sched->submit(r, id, t_err, sizeof(t_err));
}

extern "C" int grayhip_server_step(grayhip_server *v, int *busy)
{
	if (!v)
		return vfail(nullptr, GRAYHIP_EINVAL, "null server");
	return v->sched->step(busy, t_err, sizeof(t_err));
}

extern "C" int grayhip_server_next_done(grayhip_server *v, int64_t *id)
{
	if (!v || !id)
		return vfail(v, GRAYHIP_EINVAL, "null argument");
	return v->sched->next_done(id);
}

extern "C" int grayhip_server_state(grayhip_server *v, int64_t id, int *state)
{
	if (!v || !state)
		return vfail(v, GRAYHIP_EINVAL, "null argument");
	int rc = v->sched->state(id, state);
	if (rc)
		return vfail(v, rc, "unknown request %lld", (long long)id);
	return GRAYHIP_OK;
}

extern "C" int grayhip_server_result(grayhip_server *v, int64_t id, double *data, double *info,
                                     int64_t *steps)
{
	if (!v)
		return vfail(nullptr, GRAYHIP_EINVAL, "null server");
	return v->sched->result(id, data, info, steps, t_err, sizeof(t_err));
}

extern "C" int grayhip_server_stats(grayhip_server *v, int64_t steps[3], int64_t *epochs)
{
	if (!v)
		return vfail(nullptr, GRAYHIP_EINVAL, "null server");
	v->sched->stats(steps, epochs);
	return GRAYHIP_OK;
}
