/*
 * srv_mock_main.cpp -- the serving pool's host scheduler (srv_sched.cpp)
 * driven through a CPU mock of its device, for the thread and address
 * sanitizers.  TEST INFRASTRUCTURE: built and run by
 * tests/test_srv_sched_cpu.py.
 *
 * The mock follows the kernels' rules (kernels_serve.hip): a Kerr class's
 * lanes claim rays in time order (resumed rays first, then fresh ones),
 * each lane runs at most K iterations and saves the ray it still holds; a
 * Kerr ray may hand over to the FP64 list part-way; FP64 lanes claim list
 * entries the same way; grid waves claim tiles (resumed records first) and
 * run each tile until its 64 rays are done or K is used up.  Resumed
 * entries nobody claimed are carried over unchanged, and every list is
 * written from the back of its region.  Every ray takes a number of steps
 * fixed by its request tag and index, and its output row encodes tag, ray
 * and steps, so the checker knows every value in advance.
 *
 * Threads: one submits requests over time, one steps the pool, one
 * collects results (and polls state and stats) -- the concurrency the C
 * ABI allows.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <queue>
#include <thread>
#include <vector>

#include "srv_sched.h"

using namespace gray;
using namespace gray::srv;

namespace {

uint64_t mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

/* steps of ray g of the request tagged `tag`; a Kerr ray with bit 7 set
 * hands over to FP64 half-way */
int64_t ray_steps(int tag, int g) { return 20 + (int64_t)(mix64((uint64_t)tag * 1000003u + (uint64_t)g) % 2000); }
bool ray_handover(int tag, int g) { return (mix64((uint64_t)tag * 7919u + (uint64_t)g) >> 7) & 1; }

struct Ray {
	int tab, ray;
	int64_t done;        /* steps taken                       */
	bool fb;             /* in FP64                           */
};

struct Tile {
	int tab, tile;
	Ray lane[64];
	bool live[64];
};

class Mock : public Device {
public:
	Caps caps;
	std::vector<Req> table;
	std::vector<unsigned> remaining;
	std::vector<double> arena;
	std::vector<Ray> res[2], fb[2];
	std::vector<Tile> tiles[2];
	int epochs = 0;
	char err[128] = "";

	explicit Mock(const Caps &c) : caps(c), table((size_t)c.max_requests),
	                              remaining((size_t)c.max_requests, 0), arena(c.arena / 8, 0.0)
	{
		for (int k = 0; k < 2; ++k) {
			res[k].resize((size_t)c.res_cap);
			fb[k].resize((size_t)c.fb_cap);
			tiles[k].resize((size_t)c.tile_cap);
		}
	}
	const char *error() const override { return err; }

	int admit(int tab, const Req &r) override
	{
		table[(size_t)tab] = r;
		remaining[(size_t)tab] = (unsigned)r.n_rays;
		return 0;
	}

	int tag(int tab) const { return (int)table[(size_t)tab].q.i_obs; }

	void finish(const Ray &y, Result &out)
	{
		const Req &r = table[(size_t)y.tab];
		double *o = &arena[r.off / 8] + (size_t)y.ray * (size_t)r.nd;
		for (int c = 0; c < r.nd; ++c)
			o[c] = c == 0 ? (double)tag(y.tab) : (c == 1 ? (double)y.ray : (double)(c * 1000 + y.ray));
		double *info = &arena[r.off / 8] + r.n_rays * (size_t)r.nd;
		info[y.ray] = -(double)y.ray;
		int64_t *st = (int64_t *)(info + r.n_rays);
		st[y.ray] = y.done;
		if (remaining[(size_t)y.tab] == 0) {
			snprintf(err, sizeof(err), "ray finished twice");
			bad = true;
			return;
		}
		if (--remaining[(size_t)y.tab] == 0)
			out.done.push_back(y.tab);
	}

	bool bad = false;

	/* slot of the next entry written into region [base, base + cap), from
	 * the back; -1 past its start */
	int slot(unsigned &counter, int base, int cap)
	{
		const unsigned k = counter++;
		if (k >= (unsigned)cap) {
			++overflow;
			return -1;
		}
		return base + cap - 1 - (int)k;
	}

	unsigned overflow = 0;
	long carried[3] = {0, 0, 0};   /* entries carried over: Kerr, FP64, grid */

	/* a lane pool: lanes claim items in time order; run(item, t) returns the
	 * time the lane is free again, or a negative value once it saved */
	template <class Claim, class Run>
	static void pool(int lanes, int64_t K, Claim claim, Run run)
	{
		typedef std::pair<int64_t, int> Ev;
		std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> pq;
		for (int l = 0; l < lanes; ++l)
			pq.push(Ev(0, l));
		while (!pq.empty()) {
			Ev e = pq.top();
			pq.pop();
			if (e.first >= K)
				continue;            /* no refill once K is used up */
			int item;
			if (!claim(&item))
				continue;            /* exhausted */
			const int64_t t = run(item, e.first);
			if (t >= 0)
				pq.push(Ev(t, e.second));
		}
	}

	int epoch(const Plan &p, Result &out) override
	{
		const int in = p.flip, o = p.flip ^ 1;
		overflow = 0;
		unsigned fb_saved = 0;
		auto fb_put = [&](const Ray &y) {
			const int e = slot(fb_saved, 0, caps.fb_cap);
			if (e >= 0)
				fb[o][(size_t)e] = y;
		};
		/* FP64 lanes */
		{
			unsigned cur = 0;
			const int lanes = p.fb_blocks * SRV_BLOCK;
			auto claim = [&](int *item) {
				const unsigned k = cur++;
				if (k >= (unsigned)p.n_fb)
					return false;
				*item = p.fb_off + (int)k;
				return true;
			};
			auto run = [&](int item, int64_t t) -> int64_t {
				Ray y = fb[in][(size_t)item];
				const int64_t total = ray_steps(tag(y.tab), y.ray);
				const int64_t need = total - y.done;
				if (t + need <= p.K_fb) {
					out.steps[1] += (unsigned long long)need;
					y.done = total;
					finish(y, out);
					return t + need;
				}
				const int64_t k = p.K_fb - t;
				out.steps[1] += (unsigned long long)k;
				y.done += k;
				fb_put(y);
				return -1;
			};
			if (p.n_fb > 0 && lanes > 0)
				pool(lanes, p.K_fb, claim, run);
			/* entries nobody claimed are carried over */
			for (unsigned k = cur; k < (unsigned)p.n_fb; ++k, ++carried[1])
				fb_put(fb[in][(size_t)(p.fb_off + (int)k)]);
		}
		/* Kerr classes */
		for (int k = 0; k < p.n_cls; ++k) {
			const ClassPlan &c = p.cls[k];
			const int lanes = (c.block1 - c.block0) * SRV_BLOCK;
			unsigned rcur = 0, fcur = 0, saved = 0;
			auto claim = [&](int *item) {
				if (rcur < (unsigned)c.n_resume) {
					*item = c.res_off + (int)rcur++;
					return true;
				}
				const unsigned f = fcur++;
				if (f >= c.n_fresh)
					return false;
				*item = -1 - (int)f;   /* fresh ray f */
				return true;
			};
			auto run = [&](int item, int64_t t) -> int64_t {
				Ray y;
				if (item >= 0) {
					y = res[in][(size_t)item];
				} else {
					unsigned f = (unsigned)(-1 - item);
					int j = c.seg0;
					while (f >= p.kseg[(size_t)j].count) {
						f -= p.kseg[(size_t)j].count;
						++j;
					}
					y = Ray{p.kseg[(size_t)j].tab, p.kseg[(size_t)j].first + (int)f, 0, false};
				}
				const int tg = tag(y.tab);
				const int64_t total = ray_steps(tg, y.ray);
				const int64_t stop = ray_handover(tg, y.ray) && y.done < total / 2 ? total / 2 : total;
				const int64_t need = stop - y.done;
				if (t + need <= p.K_kerr) {
					out.steps[0] += (unsigned long long)need;
					y.done = stop;
					if (stop < total) {
						y.fb = true;
						fb_put(y);
					} else {
						finish(y, out);
					}
					return t + need;
				}
				const int64_t run_k = p.K_kerr - t;
				out.steps[0] += (unsigned long long)run_k;
				y.done += run_k;
				const int e = slot(saved, c.out_base, c.out_cap);
				if (e >= 0)
					res[o][(size_t)e] = y;
				return -1;
			};
			pool(lanes, p.K_kerr, claim, run);
			/* resumed rays nobody claimed are carried over */
			for (; rcur < (unsigned)c.n_resume; ++rcur, ++carried[0]) {
				const int e = slot(saved, c.out_base, c.out_cap);
				if (e >= 0)
					res[o][(size_t)e] = res[in][(size_t)(c.res_off + (int)rcur)];
			}
			out.res_claimed[k] = rcur;
			out.fresh_claimed[k] = fcur;
			out.saved[k] = saved;
		}
		/* grid waves */
		if (p.grid_blocks > 0) {
			const int waves = p.grid_blocks * SRV_GWAVES;
			unsigned rcur = 0, fcur = 0, saved = 0;
			auto claim = [&](int *item) {
				if (rcur < (unsigned)p.g_resume) {
					*item = p.g_res_off + (int)rcur++;
					return true;
				}
				const unsigned f = fcur++;
				if (f >= p.g_fresh)
					return false;
				*item = -1 - (int)f;
				return true;
			};
			auto run = [&](int item, int64_t t) -> int64_t {
				Tile T;
				if (item >= 0) {
					T = tiles[in][(size_t)item];
				} else {
					unsigned f = (unsigned)(-1 - item);
					size_t j = 0;
					while (f >= p.gseg[j].count) {
						f -= p.gseg[j].count;
						++j;
					}
					T.tab = p.gseg[j].tab;
					T.tile = p.gseg[j].first + (int)f;
					const Req &r = table[(size_t)T.tab];
					const int tw = (int)((r.q.w_rays + 7) / 8);
					for (int l = 0; l < 64; ++l) {
						const int x = (T.tile % tw) * 8 + (l & 7), yy = (T.tile / tw) * 8 + (l >> 3);
						T.live[l] = x < (int)r.q.w_rays && yy < (int)r.q.h_rays;
						T.lane[l] = Ray{T.tab, T.live[l] ? yy * (int)r.q.w_rays + x : -1, 0, false};
					}
				}
				int64_t longest = 0;
				for (int l = 0; l < 64; ++l)
					if (T.live[l])
						longest = std::max(longest, ray_steps(tag(T.tab), T.lane[l].ray) -
						                                T.lane[l].done);
				const int64_t run_k = std::min<int64_t>(longest, p.K_grid - t);
				bool any = false;
				for (int l = 0; l < 64; ++l) {
					if (!T.live[l])
						continue;
					Ray &y = T.lane[l];
					const int64_t left = ray_steps(tag(T.tab), y.ray) - y.done;
					const int64_t k2 = std::min(left, run_k);
					out.steps[2] += (unsigned long long)k2;
					y.done += k2;
					if (k2 == left) {
						finish(y, out);
						T.live[l] = false;
					} else {
						any = true;
					}
				}
				if (!any)
					return t + run_k;
				const int e = slot(saved, 0, p.g_out_cap);   /* K used up */
				if (e >= 0)
					tiles[o][(size_t)e] = T;
				return -1;
			};
			pool(waves, p.K_grid, claim, run);
			for (; rcur < (unsigned)p.g_resume; ++rcur, ++carried[2]) {
				const int e = slot(saved, 0, p.g_out_cap);
				if (e >= 0)
					tiles[o][(size_t)e] = tiles[in][(size_t)(p.g_res_off + (int)rcur)];
			}
			out.g_res = rcur;
			out.g_fresh = fcur;
			out.g_saved = saved;
		}
		out.fb_saved = fb_saved;
		out.overflow = overflow;
		out.ms[0] = out.ms[1] = out.ms[2] = -1.0;
		++epochs;
		return bad ? GRAYHIP_EDEVICE : 0;
	}

	int fetch(const Req &r, double *data, double *info, int64_t *steps) override
	{
		const double *base = &arena[r.off / 8];
		if (data)
			memcpy(data, base, r.n_rays * (size_t)r.nd * 8);
		if (info)
			memcpy(info, base + r.n_rays * (size_t)r.nd, r.n_rays * 8);
		if (steps)
			memcpy(steps, base + r.n_rays * (size_t)(r.nd + 1), r.n_rays * 8);
		return 0;
	}
};

struct Item {
	int tag, kind, w, h, nf;
};

Req make_req(const Item &it)
{
	Req r;
	memset(&r, 0, sizeof(r));
	r.q.w_rays = (size_t)it.w;
	r.q.h_rays = (size_t)it.h;
	r.q.n_freq = it.nf;
	r.q.i_obs = it.tag;
	r.kind = it.kind;
	r.nd = 8 + 2 * it.nf;
	r.n_rays = (size_t)it.w * (size_t)it.h;
	r.units = it.kind == SRV_KERR ? r.n_rays : (size_t)((it.w + 7) / 8) * (size_t)((it.h + 7) / 8);
	r.bytes = (r.n_rays * (size_t)(r.nd + 2) * 8 + 255) & ~(size_t)255;
	if (it.kind == SRV_KERR) {
		r.key.dt = -32.0;
		r.key.n_sub = 1024;
		r.key.n_dump = 128;
		r.key.n_freq = it.nf;
		/* a few distinct classes, more than fit at once */
		r.key.r_esc = 1000.0 + (it.tag % 11);
	}
	return r;
}

bool check(const Item &it, const std::vector<double> &d, const std::vector<double> &info,
           const std::vector<int64_t> &st)
{
	const int nd = 8 + 2 * it.nf, n = it.w * it.h;
	for (int g = 0; g < n; ++g) {
		if (d[(size_t)g * nd] != it.tag || d[(size_t)g * nd + 1] != g || info[(size_t)g] != -g ||
		    st[(size_t)g] != ray_steps(it.tag, g))
			return false;
		for (int c = 2; c < nd; ++c)
			if (d[(size_t)g * nd + c] != c * 1000 + g)
				return false;
	}
	return true;
}

} // namespace

int main()
{
	Caps caps;
	caps.kerr_blocks = 6;
	caps.grid_blocks = 4;
	caps.n_cu = 2;
	caps.kerr_beside = 1;
	caps.res_cap = 2 * (caps.kerr_blocks + SRV_MAX_CLASSES) * SRV_BLOCK;
	caps.tile_cap = 2 * (caps.grid_blocks * SRV_GWAVES + SRV_GWAVES);
	caps.fb_cap = 1 << 14;
	caps.max_segs = 1 << 12;
	caps.max_requests = 24;
	caps.arena = 3u << 20;
	Mock mock(caps);
	Config cfg;
	cfg.epoch_steps = 240;
	cfg.epoch_ms = 0.0;
	Sched sched(cfg, caps, &mock);

	const int N = 240;
	std::vector<Item> items((size_t)N);
	uint64_t s = 0x47524159;
	for (int k = 0; k < N; ++k) {
		s = mix64(s + 1);
		Item &it = items[(size_t)k];
		it.tag = 1000 + k;
		it.kind = (s & 3) == 0 ? SRV_GRID : SRV_KERR;
		it.w = 1 + (int)((s >> 8) % 60);
		it.h = 1 + (int)((s >> 20) % 40);
		it.nf = (int)((s >> 32) % 5);
	}
	std::vector<int64_t> ids((size_t)N, -1);
	std::vector<std::atomic<int>> seen((size_t)N);
	for (auto &a : seen)
		a = 0;
	std::atomic<int> submitted(0), checked(0), failures(0);
	std::atomic<bool> stop(false);

	std::thread feeder([&] {
		for (int k = 0; k < N; ++k) {
			int64_t id = -1;
			char e[256];
			if (sched.submit(make_req(items[(size_t)k]), &id, e, sizeof(e))) {
				fprintf(stderr, "submit: %s\n", e);
				++failures;
			}
			ids[(size_t)k] = id;
			submitted.store(k + 1);
			if (k % 16 == 0)
				std::this_thread::sleep_for(std::chrono::microseconds(200));
		}
	});
	std::thread collector([&] {
		while (!stop.load()) {
			int64_t id;
			if (sched.next_done(&id) != 0) {
				int64_t st3[3], ep;
				sched.stats(st3, &ep);
				std::this_thread::yield();
				continue;
			}
			/* ids are issued in submission order from 1 */
			const int k = (int)(id - 1);
			if (k < 0 || k >= N) {
				++failures;
				continue;
			}
			int state = -1;
			if (sched.state(id, &state) || state != GRAYHIP_REQ_DONE)
				++failures;
			const Item &it = items[(size_t)k];
			const size_t n = (size_t)it.w * (size_t)it.h;
			std::vector<double> d(n * (size_t)(8 + 2 * it.nf)), info(n);
			std::vector<int64_t> st(n);
			char e[256];
			if (sched.result(id, d.data(), info.data(), st.data(), e, sizeof(e))) {
				fprintf(stderr, "result: %s\n", e);
				++failures;
				continue;
			}
			if (!check(it, d, info, st)) {
				fprintf(stderr, "request %d: wrong output\n", k);
				++failures;
			}
			if (seen[(size_t)k]++)
				++failures;
			++checked;
		}
	});
	int epochs = 0;
	for (;;) {
		int busy = 0;
		char e[256];
		if (sched.step(&busy, e, sizeof(e))) {
			fprintf(stderr, "step: %s\n", e);
			++failures;
			break;
		}
		++epochs;
		if (!busy && submitted.load() == N) {
			/* all submitted; drained once the collector has every result */
			int64_t st3[3], ep;
			sched.stats(st3, &ep);
			if (checked.load() == N)
				break;
			std::this_thread::yield();
		}
		if (epochs > 200000) {
			fprintf(stderr, "did not drain\n");
			++failures;
			break;
		}
	}
	feeder.join();
	stop.store(true);
	collector.join();
	/* the API's error answers */
	{
		char e[256];
		double d;
		int st;
		if (sched.result(123456789, &d, nullptr, nullptr, e, sizeof(e)) != GRAYHIP_EINVAL ||
		    sched.state(123456789, &st) != GRAYHIP_EINVAL)
			++failures;
		Req big = make_req(Item{1, SRV_KERR, 4096, 4096, 4});
		int64_t id;
		if (sched.submit(big, &id, e, sizeof(e)) != GRAYHIP_EINVAL)
			++failures;
	}
	int64_t st3[3], ep;
	sched.stats(st3, &ep);
	printf("requests %d checked %d epochs %lld steps %lld %lld %lld failures %d\n", N,
	       checked.load(), (long long)ep, (long long)st3[0], (long long)st3[1], (long long)st3[2],
	       failures.load());
	if (failures.load() || checked.load() != N || st3[1] == 0 || st3[2] == 0)
		return 1;
	printf("carried kerr %ld fp64 %ld grid %ld\n", mock.carried[0], mock.carried[1],
	       mock.carried[2]);
	printf("OK\n");
	return 0;
}
