/*
 * srv_sched.cpp -- host scheduler of the serving pool (srv_sched.h).
 */
#include "srv_sched.h"

#include <algorithm>
#include <cstdio>
#include <cstring>

namespace gray {
namespace srv {

namespace {

/* Priors of the work model that shares the GPU between the Kerr and grid
 * launches, until measurements take over: steps per unit (a Kerr ray; a
 * grid tile of 64 rays of ~1e4 steps) and block-milliseconds per step (a
 * lane takes ~4.4 us per Kerr step, ~20 us per gridded step with
 * channels; 256 lanes per block). */
constexpr double PRIOR_SPU[2] = {6.4e4, 64.0 * 1e4};
constexpr double PRIOR_BMS[2] = {4.4e-3 / SRV_BLOCK, 2.0e-2 / SRV_BLOCK};
constexpr double PRIOR_UNITS[2] = {4096.0, 64.0};   /* weight of the priors */
constexpr size_t KERR_CHUNK = 256;   /* rays per request per queue round   */
constexpr size_t GRID_CHUNK = 4;     /* tiles per request per queue round  */
constexpr size_t GRID_FRESH = 4;     /* fresh tiles offered per grid wave  */
constexpr int GRID_K_BOUND = 1 << 20;

void set_err(char *err, size_t len, const char *msg)
{
	if (err && len)
		snprintf(err, len, "%s", msg);
}

} // namespace

bool ClassKey::operator==(const ClassKey &o) const
{
	if (dt != o.dt || n_sub != o.n_sub || n_dump != o.n_dump || n_freq != o.n_freq ||
	    fluid != o.fluid || escape != o.escape || r_esc != o.r_esc)
		return false;
	for (int i = 0; i < SRV_NF; ++i)
		if (nu[i] != o.nu[i])
			return false;
	return true;
}

void Sched::Arena::init(size_t size)
{
	free_.clear();
	free_[0] = size;
}

bool Sched::Arena::alloc(size_t bytes, size_t *off)
{
	for (auto it = free_.begin(); it != free_.end(); ++it)
		if (it->second >= bytes) {
			*off = it->first;
			const size_t rest = it->second - bytes, noff = it->first + bytes;
			free_.erase(it);
			if (rest)
				free_[noff] = rest;
			return true;
		}
	return false;
}

void Sched::Arena::release(size_t off, size_t bytes)
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

Sched::Sched(const Config &cfg, const Caps &caps, Device *dev)
    : cfg_(cfg), caps_(caps), dev_(dev), g_resume_(0), n_fb_(0), flip_(0), next_id_(1),
      epochs_(0)
{
	arena_.init(caps.arena);
	for (int i = caps.max_requests - 1; i >= 0; --i)
		free_tab_.push_back(i);
	tab_req_.assign((size_t)caps.max_requests, -1);
	memset(cls_, 0, sizeof(cls_));
	for (int c = 0; c < SRV_MAX_CLASSES; ++c)
		cls_used_[c] = false;
	const int K = cfg.epoch_steps > 0 ? cfg.epoch_steps : 2048;
	K_[0] = K;
	K_[1] = K;
	K_[2] = std::max(16, K / 4);
	steps_[0] = steps_[1] = steps_[2] = 0;
	for (int k = 0; k < 2; ++k) {
		units_[k] = 0.0;
		bms_[k] = PRIOR_BMS[k];
		bms_seen_[k] = false;
	}
}

/* Steps per unit of kind k (0 Kerr ray, 1 grid tile): the steps taken over
 * the units started, with the prior's weight. */
double Sched::spu(int k) const
{
	const double st = (double)(k == 0 ? steps_[0] + steps_[1] : steps_[2]);
	return (st + PRIOR_SPU[k] * PRIOR_UNITS[k]) / (units_[k] + PRIOR_UNITS[k]);
}

int Sched::submit(Req r, int64_t *id, char *err, size_t errlen)
{
	if (r.bytes > caps_.arena || r.units == 0) {
		char msg[160];
		snprintf(msg, sizeof(msg), "request needs %zu bytes of %zu, or has no rays", r.bytes,
		         caps_.arena);
		set_err(err, errlen, msg);
		return GRAYHIP_EINVAL;
	}
	std::lock_guard<std::mutex> lk(mu_);
	r.id = next_id_++;
	r.state = GRAYHIP_REQ_QUEUED;
	r.tab = -1;
	r.cls = -1;
	r.next = 0;
	r.off = 0;
	reqs_[r.id] = r;
	pending_.push_back(r.id);
	*id = r.id;
	return GRAYHIP_OK;
}

/* Admit queued requests in arrival order while a table entry, a Kerr class
 * and output memory are free. */
int Sched::admit_locked(char *err, size_t errlen)
{
	while (!pending_.empty() && !free_tab_.empty()) {
		Req &r = reqs_[pending_.front()];
		int c = -1;
		if (r.kind == SRV_KERR) {
			for (int k = 0; k < SRV_MAX_CLASSES && c < 0; ++k)
				if (cls_used_[k] && cls_[k].key == r.key)
					c = k;
			for (int k = 0; k < SRV_MAX_CLASSES && c < 0; ++k)
				if (!cls_used_[k])
					c = k;
			if (c < 0)
				break;   /* every class is taken: wait for one to empty */
		}
		size_t off;
		if (!arena_.alloc(r.bytes, &off))
			break;       /* wait until finished requests release memory */
		r.off = off;
		r.tab = free_tab_.back();
		free_tab_.pop_back();
		if (c >= 0) {
			if (!cls_used_[c]) {
				cls_used_[c] = true;
				cls_[c].key = r.key;
				cls_[c].refs = 0;
				cls_[c].n_resume = 0;
				cls_[c].res_off = 0;
			}
			++cls_[c].refs;
			r.cls = c;
		}
		int rc = dev_->admit(r.tab, r);
		if (rc) {
			set_err(err, errlen, dev_->error());
			return rc;
		}
		r.state = GRAYHIP_REQ_RUNNING;
		tab_req_[(size_t)r.tab] = r.id;
		running_.push_back(r.id);
		pending_.pop_front();
	}
	return GRAYHIP_OK;
}

/* Round-robin segments over the given requests (oldest first), `chunk`
 * units each per round, until `want` units are queued or none are left. */
void Sched::fill_segs(std::vector<int64_t> &order, size_t want, size_t chunk, size_t max_segs,
                      std::vector<Seg> &out, unsigned *total)
{
	std::vector<size_t> cur(order.size());
	for (size_t k = 0; k < order.size(); ++k)
		cur[k] = reqs_[order[k]].next;
	size_t got = 0;
	bool progress = true;
	while (got < want && progress && out.size() < max_segs) {
		progress = false;
		for (size_t k = 0; k < order.size() && got < want && out.size() < max_segs; ++k) {
			const Req &r = reqs_[order[k]];
			if (cur[k] >= r.units)
				continue;
			const size_t n = std::min(chunk, r.units - cur[k]);
			out.push_back(Seg{r.tab, (int)cur[k], (unsigned)n});
			cur[k] += n;
			got += n;
			progress = true;
		}
	}
	*total = (unsigned)got;
}

void Sched::plan_locked(Plan &p)
{
	p.n_cls = 0;
	p.kseg.clear();
	p.gseg.clear();
	p.kerr_blocks = p.grid_blocks = 0;
	p.g_resume = 0;
	p.g_items = 0;
	p.flip = flip_;
	p.K_kerr = K_[0];
	p.K_fb = K_[1];
	/* the grid launch stops on time rather than after a count of
	 * iterations: its waves' iterations differ in cost by the cells they
	 * read, and a count would leave the fast ones idle while the slowest
	 * finishes (K_grid stays as a bound) */
	p.grid_ms = cfg_.epoch_ms;
	p.K_grid = cfg_.epoch_ms > 0 ? GRID_K_BOUND : K_[2];
	p.n_fb = n_fb_;

	/* work left per class and for the grid kind */
	double wk[SRV_MAX_CLASSES] = {0};
	std::vector<int64_t> members[SRV_MAX_CLASSES], grid_reqs;
	double wg = 0.0;
	for (int64_t id : running_) {
		const Req &r = reqs_[id];
		const double left = (double)(r.units - r.next);
		if (r.kind == SRV_KERR) {
			wk[r.cls] += left;
			if (r.next < r.units)
				members[r.cls].push_back(id);
		} else {
			wg += left;
			if (r.next < r.units)
				grid_reqs.push_back(id);
		}
	}
	double wk_all = 0.0;
	int act[SRV_MAX_CLASSES], n_act = 0;
	for (int c = 0; c < SRV_MAX_CLASSES; ++c) {
		if (!cls_used_[c])
			continue;
		wk[c] += cls_[c].n_resume;
		if (wk[c] > 0) {
			act[n_act++] = c;
			wk_all += wk[c];
		}
	}
	const double tiles_left = wg + g_resume_;

	/* the GPU shared between the two launches by the time their work left
	 * would take on all of it: units x steps per unit x block-ms per step /
	 * blocks the GPU holds */
	int kerr_total = caps_.kerr_blocks, grid_total = caps_.grid_blocks;
	if (wk_all > 0 && tiles_left > 0) {
		const double tk = wk_all * spu(0) * bms_[0] / caps_.kerr_blocks;
		const double tg = tiles_left * spu(1) * bms_[1] / caps_.grid_blocks;
		const double fg = tg / (tg + tk);
		kerr_total = std::max(1, (int)(caps_.kerr_blocks * (1.0 - fg) + 0.5));
		grid_total = std::max(1, (int)(caps_.grid_blocks * fg + 0.5));
	}

	/* Kerr: every class at least the blocks its resumed rays need (all of
	 * them are claimed on the first refill), the rest by work, no more
	 * than its work can fill */
	if (n_act > 0) {
		int b[SRV_MAX_CLASSES], want[SRV_MAX_CLASSES], used = 0;
		for (int k = 0; k < n_act; ++k) {
			const int c = act[k];
			const int mn = std::max(1, (cls_[c].n_resume + SRV_BLOCK - 1) / SRV_BLOCK);
			want[k] = std::max(mn, (int)((wk[c] + SRV_BLOCK - 1) / SRV_BLOCK));
			b[k] = mn;
			used += mn;
		}
		int rest = kerr_total - used;
		if (rest > 0) {
			double w_open = 0.0;
			for (int k = 0; k < n_act; ++k)
				if (b[k] < want[k])
					w_open += wk[act[k]];
			int given = 0;
			for (int k = 0; k < n_act && w_open > 0; ++k) {
				if (b[k] >= want[k])
					continue;
				const int add = std::min(want[k] - b[k], (int)(rest * (wk[act[k]] / w_open)));
				b[k] += add;
				given += add;
			}
			rest -= given;
			for (int k = 0; k < n_act && rest > 0; ++k) {
				const int add = std::min(want[k] - b[k], rest);
				b[k] += add;
				rest -= add;
			}
		}
		/* fresh rays: twice the class's lanes, within the FP64 list's
		 * capacity (every ray of the epoch may be handed over once) */
		size_t fb_room = caps_.fb_cap > n_fb_ ? (size_t)(caps_.fb_cap - n_fb_) : 0;
		int block = 0;
		for (int k = 0; k < n_act; ++k) {
			const int c = act[k];
			ClassPlan &cp = p.cls[p.n_cls];
			cp.key = cls_[c].key;
			cp.block0 = block;
			cp.block1 = block + b[k];
			block = cp.block1;
			cp.res_off = cls_[c].res_off;
			cp.n_resume = cls_[c].n_resume;
			fb_room = fb_room > (size_t)cp.n_resume ? fb_room - (size_t)cp.n_resume : 0;
			size_t want_fresh = 2 * (size_t)b[k] * SRV_BLOCK;
			want_fresh = want_fresh > (size_t)cp.n_resume ? want_fresh - (size_t)cp.n_resume : 0;
			want_fresh = std::min(want_fresh, fb_room);
			cp.seg0 = (int)p.kseg.size();
			unsigned fresh = 0;
			fill_segs(members[c], want_fresh, KERR_CHUNK, (size_t)caps_.max_segs, p.kseg, &fresh);
			cp.seg1 = (int)p.kseg.size();
			fb_room -= fresh;
			cp.items = (unsigned)cp.n_resume + fresh;
			cp.slot = c;
			++p.n_cls;
		}
		p.kerr_blocks = block;
	}

	/* grid: at least a wave per resumed tile record, no more blocks than
	 * the tiles can fill */
	if (tiles_left > 0) {
		const int mn = std::max(1, (g_resume_ + SRV_GWAVES - 1) / SRV_GWAVES);
		const int want = std::max(mn, ((int)tiles_left + SRV_GWAVES - 1) / SRV_GWAVES);
		p.grid_blocks = std::max(mn, std::min(want, grid_total));
		p.g_resume = g_resume_;
		size_t want_fresh = GRID_FRESH * (size_t)p.grid_blocks * SRV_GWAVES;
		want_fresh = want_fresh > (size_t)g_resume_ ? want_fresh - (size_t)g_resume_ : 0;
		unsigned fresh = 0;
		fill_segs(grid_reqs, want_fresh, GRID_CHUNK, (size_t)caps_.max_segs, p.gseg, &fresh);
		p.g_items = (unsigned)g_resume_ + fresh;
	}
}

void Sched::consume_locked(const Plan &p, const Result &res)
{
	/* fresh units the queues handed out, in segment order */
	auto advance = [this](const std::vector<Seg> &segs, int s0, int s1, unsigned n) {
		for (int j = s0; j < s1 && n > 0; ++j) {
			const unsigned k = std::min(n, segs[(size_t)j].count);
			reqs_[tab_req_[(size_t)segs[(size_t)j].tab]].next += k;
			n -= k;
		}
	};
	double fresh[2] = {0.0, 0.0};
	for (int k = 0; k < p.n_cls; ++k) {
		const ClassPlan &cp = p.cls[k];
		const int c = cp.slot;
		const unsigned taken = std::min(res.claimed[k], cp.items);
		advance(p.kseg, cp.seg0, cp.seg1, taken - (unsigned)cp.n_resume);
		fresh[0] += taken - (unsigned)cp.n_resume;
		cls_[c].n_resume = (int)res.saved[k];
		cls_[c].res_off = cp.block0 * SRV_BLOCK;
	}
	/* classes that sat out this epoch hold no saved rays */
	for (int c = 0; c < SRV_MAX_CLASSES; ++c) {
		bool in = false;
		for (int k = 0; k < p.n_cls; ++k)
			in = in || p.cls[k].slot == c;
		if (!in)
			cls_[c].n_resume = 0;
	}
	if (p.g_items > 0) {
		const unsigned taken = std::min(res.g_claimed, p.g_items);
		advance(p.gseg, 0, (int)p.gseg.size(), taken - (unsigned)p.g_resume);
		fresh[1] += taken - (unsigned)p.g_resume;
		g_resume_ = (int)res.g_saved;
	} else {
		g_resume_ = 0;
	}
	n_fb_ = (int)res.fb_out;
	flip_ ^= 1;
	for (int tab : res.done) {
		const int64_t id = tab_req_[(size_t)tab];
		Req &r = reqs_[id];
		r.state = GRAYHIP_REQ_DONE;
		tab_req_[(size_t)tab] = -1;
		free_tab_.push_back(tab);
		if (r.kind == SRV_KERR && --cls_[r.cls].refs == 0)
			cls_used_[r.cls] = false;
		running_.erase(std::find(running_.begin(), running_.end(), id));
		done_.push_back(id);
	}
	for (int k = 0; k < 3; ++k)
		steps_[k] += res.steps[k];
	++epochs_;
	/* work model: units started and steps taken per kind; block-ms per
	 * step from epochs whose launch was busy throughout (it saved work) */
	units_[0] += fresh[0];
	units_[1] += fresh[1];
	const double ksteps[2] = {(double)res.steps[0], (double)res.steps[2]};
	const double kms[2] = {res.ms[0], res.ms[2]};
	const int kblk[2] = {p.kerr_blocks, p.grid_blocks};
	bool kfull[2] = {false, res.g_saved > 0};
	for (int k = 0; k < p.n_cls; ++k)
		kfull[0] = kfull[0] || res.saved[k] > 0;
	for (int k = 0; k < 2; ++k)
		if (kfull[k] && kms[k] > 0.0 && ksteps[k] > 0.0) {
			const double b = kms[k] * kblk[k] / ksteps[k];
			bms_[k] = bms_seen_[k] ? 0.7 * bms_[k] + 0.3 * b : b;
			bms_seen_[k] = true;
		}
	/* iterations per wave toward the target epoch time, from kernels that
	 * ran them all (they saved work for the next epoch) */
	if (cfg_.epoch_ms > 0) {
		bool full[3] = {false, res.fb_out > 0, res.g_saved > 0};
		for (int k = 0; k < p.n_cls; ++k)
			full[0] = full[0] || res.saved[k] > 0;
		for (int k = 0; k < 3; ++k) {
			if (!full[k] || !(res.ms[k] > 0.0))
				continue;
			double f = cfg_.epoch_ms / res.ms[k];
			f = std::min(2.0, std::max(0.5, f));
			K_[k] = std::max(16, std::min(1 << 20, (int)(K_[k] * f)));
		}
	}
}

int Sched::step(int *busy, char *err, size_t errlen)
{
	Plan p;
	{
		std::lock_guard<std::mutex> lk(mu_);
		int rc = admit_locked(err, errlen);
		if (rc)
			return rc;
		plan_locked(p);
		const bool work = p.n_cls > 0 || p.g_items > 0 || p.n_fb > 0;
		if (busy)
			*busy = work || !pending_.empty();
		if (!work)
			return GRAYHIP_OK;
	}
	Result res;
	memset(res.claimed, 0, sizeof(res.claimed));
	memset(res.saved, 0, sizeof(res.saved));
	res.fb_out = res.g_claimed = res.g_saved = 0;
	res.steps[0] = res.steps[1] = res.steps[2] = 0;
	res.ms[0] = res.ms[1] = res.ms[2] = -1.0;
	int rc = dev_->epoch(p, res);
	std::lock_guard<std::mutex> lk(mu_);
	if (rc) {
		set_err(err, errlen, dev_->error());
		return rc;
	}
	/* every resumed ray and tile is claimed before any fresh one */
	for (int k = 0; k < p.n_cls; ++k)
		if (res.claimed[k] < (unsigned)p.cls[k].n_resume) {
			set_err(err, errlen, "epoch left resumed rays unclaimed");
			return GRAYHIP_EDEVICE;
		}
	if (p.g_items > 0 && res.g_claimed < (unsigned)p.g_resume) {
		set_err(err, errlen, "epoch left resumed tiles unclaimed");
		return GRAYHIP_EDEVICE;
	}
	if (res.fb_out > (unsigned)caps_.fb_cap) {
		set_err(err, errlen, "FP64 list overflowed");
		return GRAYHIP_EDEVICE;
	}
	consume_locked(p, res);
	return GRAYHIP_OK;
}

int Sched::next_done(int64_t *id)
{
	std::lock_guard<std::mutex> lk(mu_);
	if (done_.empty())
		return 1;
	*id = done_.front();
	done_.pop_front();
	return GRAYHIP_OK;
}

int Sched::state(int64_t id, int *st)
{
	std::lock_guard<std::mutex> lk(mu_);
	auto it = reqs_.find(id);
	if (it == reqs_.end())
		return GRAYHIP_EINVAL;
	*st = it->second.state;
	return GRAYHIP_OK;
}

int Sched::result(int64_t id, double *data, double *info, int64_t *steps, char *err,
                  size_t errlen)
{
	std::lock_guard<std::mutex> lk(mu_);
	auto it = reqs_.find(id);
	if (it == reqs_.end()) {
		set_err(err, errlen, "unknown request");
		return GRAYHIP_EINVAL;
	}
	Req &r = it->second;
	if (r.state != GRAYHIP_REQ_DONE) {
		set_err(err, errlen, "request is not finished");
		return GRAYHIP_ESTATE;
	}
	int rc = dev_->fetch(r, data, info, steps);
	if (rc) {
		set_err(err, errlen, dev_->error());
		return rc;
	}
	arena_.release(r.off, r.bytes);
	done_.erase(std::remove(done_.begin(), done_.end(), id), done_.end());
	reqs_.erase(it);
	return GRAYHIP_OK;
}

void Sched::stats(int64_t steps[3], int64_t *epochs)
{
	std::lock_guard<std::mutex> lk(mu_);
	if (steps)
		for (int k = 0; k < 3; ++k)
			steps[k] = (int64_t)steps_[k];
	if (epochs)
		*epochs = epochs_;
}

} // namespace srv
} // namespace gray
