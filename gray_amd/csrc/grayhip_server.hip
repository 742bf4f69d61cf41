/*
 * grayhip_server.hip -- the serving pool (include/grayhip.h,
 * grayhip_server_*): request queue, admission by an HBM budget, slot
 * refill, and one bounded epoch per step over three slot lists (FP32 Kerr,
 * FP64 Kerr re-integration, FP32 grid) on three streams, so the FP64 work
 * runs beside the FP32 work instead of after it.
 */
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <vector>

#include "launch_serve.h"
#include "spacetime.h"

using namespace gray;

namespace {

struct HReq {
	int64_t id;
	int state;
	int tab;                 /* index in the device request table */
	int kind, nd;
	size_t n_rays, next, remaining;
	size_t off, bytes;       /* output region in the arena */
	SrvRequest dev;
};

/* First-fit allocator of arena offsets with coalescing. */
struct Arena {
	std::map<size_t, size_t> free_; /* offset -> size */
	void init(size_t size) { free_.clear(); free_[0] = size; }
	bool alloc(size_t bytes, size_t *off)
	{
		for (auto it = free_.begin(); it != free_.end(); ++it)
			if (it->second >= bytes) {
				*off = it->first;
				size_t rest = it->second - bytes, noff = it->first + bytes;
				free_.erase(it);
				if (rest)
					free_[noff] = rest;
				return true;
			}
		return false;
	}
	void release(size_t off, size_t bytes)
	{
		auto it = free_.emplace(off, bytes).first;
		auto nx = std::next(it);
		if (nx != free_.end() && it->first + it->second == nx->first) {
			it->second += nx->second;
			free_.erase(nx);
		}
		if (it != free_.begin()) {
			auto pv = std::prev(it);
			if (pv->first + pv->second == it->first) {
				pv->second += it->second;
				free_.erase(it);
			}
		}
	}
};

} // namespace

struct grayhip_server {
	grayhip_server_config cfg;
	hipStream_t s[3];
	hipEvent_t ev_ready;
	SrvRequest *d_reqs;
	SrvSlots S;
	int *d_list[3];
	int *d_assign;            /* 3 x slots: slot, table index, ray */
	SrvLists L;
	char *arena;
	Arena alloc;
	/* pinned host mirrors */
	SrvCounters *h_ctr;
	int *h_fin, *h_to64, *h_assign;
	/* host bookkeeping */
	std::map<int64_t, HReq> reqs;
	std::deque<int64_t> pending, filling, done;
	std::vector<int> free_tab, free_slots;
	std::vector<int64_t> slot_req;   /* request id per slot */
	std::vector<int> slot_list, slot_pos;
	std::vector<int> list[3];
	int64_t next_id;
	int64_t epochs;
	unsigned long long steps[3];
	/* Guards the host bookkeeping (requests, queues, arena, slot lists) so
	 * submit / next_done / state / result / stats may run on other threads
	 * while one thread steps; not held while an epoch runs on the GPU. */
	std::mutex mu;
	char err[256];
};

static int vfail(grayhip_server *v, int code, const char *fmt, ...)
{
	static thread_local char g[256];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(v ? v->err : g, 256, fmt, ap);
	va_end(ap);
	return code;
}

#define SCK(v, expr)                                                                   \
	do {                                                                           \
		hipError_t e_ = (expr);                                                \
		if (e_ != hipSuccess)                                                  \
			return vfail(v, GRAYHIP_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
	} while (0)

extern "C" const char *grayhip_server_error(const grayhip_server *v)
{
	return v ? v->err : "server: null handle or creation failed";
}

extern "C" void grayhip_server_destroy(grayhip_server *v)
{
	if (!v)
		return;
	(void)hipSetDevice(v->cfg.device);
	for (int i = 0; i < 3; ++i) {
		if (v->s[i])
			(void)hipStreamDestroy(v->s[i]);
		(void)hipFree(v->d_list[i]);
	}
	if (v->ev_ready)
		(void)hipEventDestroy(v->ev_ready);
	(void)hipFree(v->d_reqs);
	(void)hipFree(v->S.req);
	(void)hipFree(v->S.ray);
	(void)hipFree(v->S.d);
	(void)hipFree(v->S.h);
	(void)hipFree(v->S.steps);
	(void)hipFree(v->S.st);
	(void)hipFree(v->S.st64);
	(void)hipFree(v->d_assign);
	(void)hipFree(v->L.fin);
	(void)hipFree(v->L.to64);
	(void)hipFree(v->L.ctr);
	(void)hipFree(v->arena);
	(void)hipHostFree(v->h_ctr);
	(void)hipHostFree(v->h_fin);
	(void)hipHostFree(v->h_to64);
	(void)hipHostFree(v->h_assign);
	delete v;
}

extern "C" int grayhip_server_create(const grayhip_server_config *cfg, grayhip_server **out)
{
	if (!cfg || !out)
		return vfail(nullptr, GRAYHIP_EINVAL, "null argument");
	*out = nullptr;
	if (cfg->slots == 0 || cfg->slots > (1u << 28) || cfg->epoch_steps <= 0 ||
	    cfg->max_requests <= 0 || cfg->hbm_budget == 0)
		return vfail(nullptr, GRAYHIP_EINVAL, "bad server configuration");
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
		return vfail(nullptr, GRAYHIP_ENODEV, "no HIP device visible");
	grayhip_server *v = new (std::nothrow) grayhip_server();
	if (!v)
		return vfail(nullptr, GRAYHIP_ENOMEM, "host allocation failed");
	v->cfg = *cfg;
	const size_t S = cfg->slots, R = (size_t)cfg->max_requests;
	hipError_t e = hipSetDevice(cfg->device);
	struct { void **p; size_t b; } al[] = {
		{(void **)&v->d_reqs, R * sizeof(SrvRequest)},
		{(void **)&v->S.req, S * sizeof(int)},
		{(void **)&v->S.ray, S * sizeof(int)},
		{(void **)&v->S.d, S * sizeof(int)},
		{(void **)&v->S.h, S * sizeof(int)},
		{(void **)&v->S.steps, S * sizeof(int64_t)},
		{(void **)&v->S.st, S * 2 * SRV_NC * sizeof(float)},
		{(void **)&v->S.st64, S * SRV_NC * sizeof(double)},
		{(void **)&v->d_list[0], S * sizeof(int)},
		{(void **)&v->d_list[1], S * sizeof(int)},
		{(void **)&v->d_list[2], S * sizeof(int)},
		{(void **)&v->d_assign, 3 * S * sizeof(int)},
		{(void **)&v->L.fin, S * sizeof(int)},
		{(void **)&v->L.to64, S * sizeof(int)},
		{(void **)&v->L.ctr, sizeof(SrvCounters)},
		{(void **)&v->arena, cfg->hbm_budget},
	};
	for (auto &a : al)
		if (e == hipSuccess)
			e = hipMalloc(a.p, a.b);
	if (e == hipSuccess)
		e = hipHostMalloc((void **)&v->h_ctr, sizeof(SrvCounters), 0);
	if (e == hipSuccess)
		e = hipHostMalloc((void **)&v->h_fin, S * sizeof(int), 0);
	if (e == hipSuccess)
		e = hipHostMalloc((void **)&v->h_to64, S * sizeof(int), 0);
	if (e == hipSuccess)
		e = hipHostMalloc((void **)&v->h_assign, 3 * S * sizeof(int), 0);
	for (int i = 0; i < 3 && e == hipSuccess; ++i)
		e = hipStreamCreateWithFlags(&v->s[i], hipStreamNonBlocking);
	if (e == hipSuccess)
		e = hipEventCreateWithFlags(&v->ev_ready, hipEventDisableTiming);
	if (e != hipSuccess) {
		int rc = vfail(nullptr, GRAYHIP_ENOMEM, "server allocation: %s", hipGetErrorString(e));
		grayhip_server_destroy(v);
		return rc;
	}
	v->S.S = (int)S;
	v->alloc.init(cfg->hbm_budget);
	for (int i = (int)R - 1; i >= 0; --i)
		v->free_tab.push_back(i);
	v->free_slots.reserve(S);
	for (int i = (int)S - 1; i >= 0; --i)
		v->free_slots.push_back(i);
	v->slot_req.assign(S, -1);
	v->slot_list.assign(S, -1);
	v->slot_pos.assign(S, -1);
	v->next_id = 1;
	*out = v;
	return GRAYHIP_OK;
}

static double r_isco_host(double a)
{
	double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
	double z2 = sqrt(3.0 * a * a + z1 * z1);
	return 3.0 + z2 - sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
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
	    q->n_dump == 0 || q->n_dump > (size_t)INT32_MAX || !(q->dt != 0.0))
		return vfail(v, GRAYHIP_EINVAL, "bad spin or schedule");
	if (q->n_freq > 0 && q->fluid != GRAYHIP_FLUID_DISK && q->fluid != GRAYHIP_FLUID_UNIFORM)
		return vfail(v, GRAYHIP_EINVAL, "served analytic requests use the disk or uniform fluid");
	if (q->spacetime && (q->spacetime->desc.dtype != 4 || q->n_freq != 0 ||
	                     q->spacetime->desc.device != v->cfg.device))
		return vfail(v, GRAYHIP_EINVAL, "gridded requests need an FP32 spacetime on this "
		                                "device and no transfer channels");
	HReq r;
	memset(&r, 0, sizeof(r));
	r.state = GRAYHIP_REQ_QUEUED;
	r.tab = -1;
	r.kind = q->spacetime ? SRV_GRID : SRV_KERR;
	r.nd = 8 + 2 * q->n_freq;
	r.n_rays = q->w_rays * q->h_rays;
	r.remaining = r.n_rays;
	r.bytes = (r.n_rays * (size_t)(r.nd + 2) * 8 + 255) & ~(size_t)255;
	if (r.bytes > v->cfg.hbm_budget)
		return vfail(v, GRAYHIP_EINVAL, "request needs %zu bytes, budget is %zu", r.bytes,
		             v->cfg.hbm_budget);
	RayParams &p = r.dev.p;
	p.a = q->spacetime ? 0.0 : q->a_spin;
	p.r_h = 1.0 + sqrt(1.0 - p.a * p.a);
	p.escape = q->escape;
	p.eps_stop = q->eps_stop > 0.0 ? q->eps_stop
	             : (q->escape == GRAYHIP_ESC_OUTWARD ? 0.01 * p.r_h : 0.01);
	p.r_esc = q->r_esc;
	p.r_isco = r_isco_host(p.a);
	p.n_freq = q->n_freq;
	p.fluid = q->n_freq > 0 ? q->fluid : GRAYHIP_FLUID_NONE;
	for (int i = 0; i < q->n_freq; ++i) {
		if (!(q->nu[i] > 0.0))
			return vfail(v, GRAYHIP_EINVAL, "frequency channel %d must be positive", i);
		p.nu[i] = q->nu[i];
	}
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
	StepRule &sr = r.dev.sr;
	sr.dt = q->dt;
	sr.dts = q->dt / (double)q->n_sub;
	sr.n = (int)((INT32_MAX / (long)q->n_sub) * (long)q->n_sub);
	sr.n_dump = (int)q->n_dump;
	if (q->spacetime)
		r.dev.gv = spacetime_view(q->spacetime, q->t_init, 0, q->escape == GRAYHIP_ESC_REF)
		               .view<float>();
	r.dev.kind = r.kind;
	r.dev.nd = r.nd;
	std::lock_guard<std::mutex> lk(v->mu);
	r.id = v->next_id++;
	v->reqs[r.id] = r;
	v->pending.push_back(r.id);
	*id = r.id;
	return GRAYHIP_OK;
}

static void list_add(grayhip_server *v, int which, int slot)
{
	v->slot_list[slot] = which;
	v->slot_pos[slot] = (int)v->list[which].size();
	v->list[which].push_back(slot);
}

static void list_remove(grayhip_server *v, int slot)
{
	int w = v->slot_list[slot], p = v->slot_pos[slot];
	std::vector<int> &L = v->list[w];
	int last = L.back();
	L[p] = last;
	v->slot_pos[last] = p;
	L.pop_back();
	v->slot_list[slot] = -1;
	v->slot_pos[slot] = -1;
}

static int harvest(grayhip_server *v)
{
	SCK(v, hipMemcpy(v->h_ctr, v->L.ctr, sizeof(SrvCounters), hipMemcpyDeviceToHost));
	const unsigned nf = v->h_ctr->n_fin, n64 = v->h_ctr->n_to64;
	if (nf > (unsigned)v->S.S || n64 > (unsigned)v->S.S)
		return vfail(v, GRAYHIP_EDEVICE, "epoch lists overflowed");
	if (nf)
		SCK(v, hipMemcpy(v->h_fin, v->L.fin, nf * sizeof(int), hipMemcpyDeviceToHost));
	if (n64)
		SCK(v, hipMemcpy(v->h_to64, v->L.to64, n64 * sizeof(int), hipMemcpyDeviceToHost));
	for (int k = 0; k < 3; ++k)
		v->steps[k] += v->h_ctr->steps[k];
	for (unsigned i = 0; i < n64; ++i) {
		int s = v->h_to64[i];
		list_remove(v, s);
		list_add(v, 1, s);
	}
	for (unsigned i = 0; i < nf; ++i) {
		int s = v->h_fin[i];
		list_remove(v, s);
		HReq &r = v->reqs[v->slot_req[s]];
		v->slot_req[s] = -1;
		v->free_slots.push_back(s);
		if (--r.remaining == 0) {
			r.state = GRAYHIP_REQ_DONE;
			v->free_tab.push_back(r.tab);
			v->done.push_back(r.id);
		}
	}
	return GRAYHIP_OK;
}

static int admit(grayhip_server *v)
{
	while (!v->pending.empty() && !v->free_tab.empty()) {
		HReq &r = v->reqs[v->pending.front()];
		size_t off;
		if (!v->alloc.alloc(r.bytes, &off))
			break; /* wait until finished requests release memory */
		r.off = off;
		r.tab = v->free_tab.back();
		v->free_tab.pop_back();
		double *base = (double *)(v->arena + off);
		r.dev.out_data = base;
		r.dev.out_info = base + r.n_rays * (size_t)r.nd;
		r.dev.out_steps = (int64_t *)(r.dev.out_info + r.n_rays);
		SCK(v, hipMemcpy(v->d_reqs + r.tab, &r.dev, sizeof(SrvRequest), hipMemcpyHostToDevice));
		r.state = GRAYHIP_REQ_RUNNING;
		v->filling.push_back(r.id);
		v->pending.pop_front();
	}
	return GRAYHIP_OK;
}

/* Give free slots to rays of admitted requests.  The free slots are shared
 * evenly among the requests still waiting for slots (the oldest take the
 * remainder), so a small request that arrives behind a large one gets its
 * slots at the next epoch instead of after the large one's last ray. */
static int refill(grayhip_server *v, int *n_assign)
{
	const size_t S = (size_t)v->S.S;
	int n = 0;
	int *as = v->h_assign, *ar = v->h_assign + S, *ay = v->h_assign + 2 * S;
	while (!v->filling.empty() && !v->free_slots.empty()) {
		const size_t n_req = v->filling.size();
		const size_t share = v->free_slots.size() / n_req;
		size_t extra = v->free_slots.size() % n_req;
		for (size_t k = 0; k < n_req && !v->free_slots.empty(); ++k) {
			HReq &r = v->reqs[v->filling[k]];
			size_t give = share + (extra > 0 ? 1 : 0);
			if (extra > 0)
				--extra;
			for (; give > 0 && r.next < r.n_rays && !v->free_slots.empty(); --give) {
				int s = v->free_slots.back();
				v->free_slots.pop_back();
				as[n] = s;
				ar[n] = r.tab;
				ay[n] = (int)r.next++;
				v->slot_req[s] = r.id;
				list_add(v, r.kind == SRV_GRID ? 2 : 0, s);
				++n;
			}
		}
		/* drop requests whose rays all have slots; slots a full request
		 * left unused go round again */
		std::deque<int64_t> still;
		for (int64_t id : v->filling)
			if (v->reqs[id].next < v->reqs[id].n_rays)
				still.push_back(id);
		v->filling.swap(still);
	}
	*n_assign = n;
	if (n) {
		int *d = v->d_assign;
		SCK(v, hipMemcpyAsync(d, as, n * sizeof(int), hipMemcpyHostToDevice, v->s[0]));
		SCK(v, hipMemcpyAsync(d + S, ar, n * sizeof(int), hipMemcpyHostToDevice, v->s[0]));
		SCK(v, hipMemcpyAsync(d + 2 * S, ay, n * sizeof(int), hipMemcpyHostToDevice, v->s[0]));
	}
	return GRAYHIP_OK;
}

extern "C" int grayhip_server_step(grayhip_server *v, int *busy)
{
	if (!v)
		return vfail(nullptr, GRAYHIP_EINVAL, "null server");
	SCK(v, hipSetDevice(v->cfg.device));
	int n_assign = 0;
	bool work;
	{
		std::lock_guard<std::mutex> lk(v->mu);
		int rc = admit(v);
		if (!rc)
			rc = refill(v, &n_assign);
		if (rc)
			return rc;
		work = !v->list[0].empty() || !v->list[1].empty() || !v->list[2].empty();
		if (busy)
			*busy = work || !v->pending.empty();
	}
	if (!work)
		return GRAYHIP_OK;
	const size_t S = (size_t)v->S.S;
	SCK(v, hipMemsetAsync(v->L.ctr, 0, sizeof(SrvCounters), v->s[0]));
	for (int k = 0; k < 3; ++k)
		if (!v->list[k].empty())
			SCK(v, hipMemcpyAsync(v->d_list[k], v->list[k].data(), v->list[k].size() * sizeof(int),
			                      hipMemcpyHostToDevice, v->s[0]));
	SCK(v, launch_srv_init(v->d_reqs, v->S, v->d_assign, v->d_assign + S, v->d_assign + 2 * S,
	                       n_assign, v->s[0]));
	SCK(v, hipEventRecord(v->ev_ready, v->s[0]));
	SCK(v, hipStreamWaitEvent(v->s[1], v->ev_ready, 0));
	SCK(v, hipStreamWaitEvent(v->s[2], v->ev_ready, 0));
	const int K = v->cfg.epoch_steps;
	SCK(v, launch_srv_epoch(0, v->d_reqs, v->S, v->d_list[0], (int)v->list[0].size(), K, v->L, v->s[0]));
	SCK(v, launch_srv_epoch(1, v->d_reqs, v->S, v->d_list[1], (int)v->list[1].size(), K, v->L, v->s[1]));
	SCK(v, launch_srv_epoch(2, v->d_reqs, v->S, v->d_list[2], (int)v->list[2].size(), K, v->L, v->s[2]));
	for (int k = 0; k < 3; ++k)
		SCK(v, hipStreamSynchronize(v->s[k]));
	std::lock_guard<std::mutex> lk(v->mu);
	++v->epochs;
	return harvest(v);
}

extern "C" int grayhip_server_next_done(grayhip_server *v, int64_t *id)
{
	if (!v || !id)
		return vfail(v, GRAYHIP_EINVAL, "null argument");
	std::lock_guard<std::mutex> lk(v->mu);
	if (v->done.empty())
		return 1;
	*id = v->done.front();
	v->done.pop_front();
	return GRAYHIP_OK;
}

extern "C" int grayhip_server_state(grayhip_server *v, int64_t id, int *state)
{
	if (!v || !state)
		return vfail(v, GRAYHIP_EINVAL, "null argument");
	std::lock_guard<std::mutex> lk(v->mu);
	auto it = v->reqs.find(id);
	if (it == v->reqs.end())
		return vfail(v, GRAYHIP_EINVAL, "unknown request %lld", (long long)id);
	*state = it->second.state;
	return GRAYHIP_OK;
}

extern "C" int grayhip_server_result(grayhip_server *v, int64_t id, double *data, double *info,
                                     int64_t *steps)
{
	if (!v)
		return vfail(nullptr, GRAYHIP_EINVAL, "null server");
	std::lock_guard<std::mutex> lk(v->mu);
	auto it = v->reqs.find(id);
	if (it == v->reqs.end())
		return vfail(v, GRAYHIP_EINVAL, "unknown request %lld", (long long)id);
	HReq &r = it->second;
	if (r.state != GRAYHIP_REQ_DONE)
		return vfail(v, GRAYHIP_ESTATE, "request %lld is not finished", (long long)id);
	SCK(v, hipSetDevice(v->cfg.device));
	if (data)
		SCK(v, hipMemcpy(data, r.dev.out_data, r.n_rays * (size_t)r.nd * 8, hipMemcpyDeviceToHost));
	if (info)
		SCK(v, hipMemcpy(info, r.dev.out_info, r.n_rays * 8, hipMemcpyDeviceToHost));
	if (steps)
		SCK(v, hipMemcpy(steps, r.dev.out_steps, r.n_rays * 8, hipMemcpyDeviceToHost));
	v->alloc.release(r.off, r.bytes);
	v->done.erase(std::remove(v->done.begin(), v->done.end(), id), v->done.end());
	v->reqs.erase(it);
	return GRAYHIP_OK;
}

extern "C" int grayhip_server_stats(grayhip_server *v, int64_t steps[3], int64_t *epochs)
{
	if (!v)
		return vfail(nullptr, GRAYHIP_EINVAL, "null server");
	std::lock_guard<std::mutex> lk(v->mu);
	if (steps)
		for (int k = 0; k < 3; ++k)
			steps[k] = (int64_t)v->steps[k];
	if (epochs)
		*epochs = v->epochs;
	return GRAYHIP_OK;
}
