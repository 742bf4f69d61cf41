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
constexpr int K_BOUND = 1 << 20;     /* iterations bound of timed launches */

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
    : cfg_(cfg), caps_(caps), dev_(dev), g_res_off_(0), g_resume_(0), fb_off_(0), n_fb_(0),
      flip_(0), next_id_(1),
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
	p.kerr_blocks = p.grid_blocks = p.fb_blocks = 0;
	p.g_res_off = g_res_off_;
	p.g_resume = g_resume_;
	p.g_out_cap = 0;
	p.g_fresh = 0;
	p.fb_off = fb_off_;
	p.n_fb = n_fb_;
	p.flip = flip_;
	/* the launches stop on time rather than after a count of iterations
	 * (their K stay as bounds): a gridded wave's iterations differ in cost
	 * by the cells it reads, and a count would leave the fast waves idle
	 * while the slowest finishes */
	p.epoch_ms = cfg_.epoch_ms;
	const bool timed = cfg_.epoch_ms > 0;
	p.K_kerr = timed ? K_BOUND : K_[0];
	p.K_fb = timed ? K_BOUND : K_[1];
	p.K_grid = timed ? K_BOUND : K_[2];

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

	/* FP64 lanes: enough for the list, on at most a quarter of the CUs (the
	 * rest is carried over; FP64 rays are few, and each of its blocks
	 * displaces Kerr ones) */
	if (n_fb_ > 0)
		p.fb_blocks = std::min(std::max(1, caps_.n_cu / 4), (n_fb_ + SRV_BLOCK - 1) / SRV_BLOCK);

	/* grid blocks: beside Kerr work, its share of the GPU by the time each
	 * kind's work left would take on all of it (units x steps per unit x
	 * block-ms per step / blocks the GPU holds), at most one FP64 or grid
	 * block per CU; alone, as many as its tiles fill */
	if (tiles_left > 0) {
		const int need = std::max(1, ((int)tiles_left + SRV_GWAVES - 1) / SRV_GWAVES);
		int G;
		if (wk_all > 0) {
			const double tk = wk_all * spu(0) * bms_[0] / caps_.kerr_blocks;
			const double tg = tiles_left * spu(1) * bms_[1] / caps_.grid_blocks;
			const double fg = tg / (tg + tk);
			const int room = std::max(1, caps_.n_cu - p.fb_blocks);
			G = std::max(1, std::min(std::min(need, room), (int)(fg * caps_.grid_blocks + 0.5)));
		} else {
			G = std::min(need, std::max(1, caps_.grid_blocks - p.fb_blocks));
		}
		p.grid_blocks = G;
		const int W = G * SRV_GWAVES;
		p.g_out_cap = std::max(W, g_resume_);
		size_t want_fresh = GRID_FRESH * (size_t)W;
		want_fresh = want_fresh > (size_t)g_resume_ ? want_fresh - (size_t)g_resume_ : 0;
		fill_segs(grid_reqs, want_fresh, GRID_CHUNK, (size_t)caps_.max_segs, p.gseg, &p.g_fresh);
	}

	/* Kerr: the room the FP64 and grid blocks leave (a CU holding one of
	 * them keeps kerr_beside Kerr blocks, one holding two none), shared
	 * among the classes by work, at least a block each */
	if (n_act > 0) {
		const int heavy = p.fb_blocks + p.grid_blocks;
		const int per_cu = std::max(1, caps_.kerr_blocks / std::max(1, caps_.n_cu));
		int total;
		if (heavy <= caps_.n_cu)
			total = caps_.kerr_blocks - (per_cu - caps_.kerr_beside) * heavy;
		else
			total = caps_.kerr_beside * std::max(0, 2 * caps_.n_cu - heavy);
		total = std::max(total, n_act);
		int b[SRV_MAX_CLASSES], want[SRV_MAX_CLASSES];
		for (int k = 0; k < n_act; ++k) {
			want[k] = std::max(1, (int)((wk[act[k]] + SRV_BLOCK - 1) / SRV_BLOCK));
			b[k] = 1;
		}
		int rest = total - n_act;
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
		 * capacity (every ray of the epoch may be handed over once) and
		 * within the rays the GPU holds (the rays in flight after the
		 * epoch, resumed or fresh, then fit the resume lists however the
		 * classes' blocks change) */
		size_t fb_room = caps_.fb_cap > n_fb_ ? (size_t)(caps_.fb_cap - n_fb_) : 0;
		size_t held = 0;
		for (int k = 0; k < n_act; ++k)
			held += (size_t)cls_[act[k]].n_resume;
		const size_t lanes_all = (size_t)caps_.kerr_blocks * SRV_BLOCK;
		size_t flight_room = lanes_all > held ? lanes_all - held : 0;
		int block = 0, out = 0;
		for (int k = 0; k < n_act; ++k) {
			const int c = act[k];
			ClassPlan &cp = p.cls[p.n_cls];
			cp.key = cls_[c].key;
			cp.slot = c;
			cp.block0 = block;
			cp.block1 = block + b[k];
			block = cp.block1;
			cp.res_off = cls_[c].res_off;
			cp.n_resume = cls_[c].n_resume;
			const int lanes = b[k] * SRV_BLOCK;
			cp.out_base = out;
			cp.out_cap = std::max(lanes, cp.n_resume);
			out += cp.out_cap;
			fb_room = fb_room > (size_t)cp.n_resume ? fb_room - (size_t)cp.n_resume : 0;
			size_t want_fresh = 2 * (size_t)lanes;
			want_fresh = want_fresh > (size_t)cp.n_resume ? want_fresh - (size_t)cp.n_resume : 0;
			want_fresh = std::min(want_fresh, std::min(fb_room, flight_room));
			cp.seg0 = (int)p.kseg.size();
			fill_segs(members[c], want_fresh, KERR_CHUNK, (size_t)caps_.max_segs, p.kseg, &cp.n_fresh);
			cp.seg1 = (int)p.kseg.size();
			fb_room -= cp.n_fresh;
			flight_room -= cp.n_fresh;
			++p.n_cls;
		}
		p.kerr_blocks = block;
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
	for (int c = 0; c < SRV_MAX_CLASSES; ++c)
		cls_[c].n_resume = 0;   /* a class with saved rays is always planned */
	for (int k = 0; k < p.n_cls; ++k) {
		const ClassPlan &cp = p.cls[k];
		const int c = cp.slot;
		const unsigned taken = std::min(res.fresh_claimed[k], cp.n_fresh);
		advance(p.kseg, cp.seg0, cp.seg1, taken);
		fresh[0] += taken;
		cls_[c].n_resume = (int)std::min(res.saved[k], (unsigned)cp.out_cap);
		cls_[c].res_off = cp.out_base + cp.out_cap - cls_[c].n_resume;
	}
	if (p.grid_blocks > 0) {
		const unsigned taken = std::min(res.g_fresh, p.g_fresh);
		advance(p.gseg, 0, (int)p.gseg.size(), taken);
		fresh[1] += taken;
		g_resume_ = (int)std::min(res.g_saved, (unsigned)p.g_out_cap);
		g_res_off_ = p.g_out_cap - g_resume_;
	} else {
		g_resume_ = g_res_off_ = 0;
	}
	n_fb_ = (int)std::min(res.fb_saved, (unsigned)caps_.fb_cap);
	fb_off_ = caps_.fb_cap - n_fb_;
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
	 * step from epochs whose launch was busy throughout (it saved work),
	 * over the blocks that started before the stop time */
	units_[0] += fresh[0];
	units_[1] += fresh[1];
	const double ksteps[2] = {(double)res.steps[0], (double)res.steps[2]};
	const double kms[2] = {res.ms[0], res.ms[2]};
	const double kblk[2] = {p.kerr_blocks - res.late[0] / 4.0, p.grid_blocks - res.late[2] / 4.0};
	bool kfull[2] = {false, res.g_saved > 0};
	for (int k = 0; k < p.n_cls; ++k)
		kfull[0] = kfull[0] || res.saved[k] > 0;
	for (int k = 0; k < 2; ++k)
		if (kfull[k] && kms[k] > 0.0 && ksteps[k] > 0.0 && kblk[k] > 0.0) {
			const double b = kms[k] * kblk[k] / ksteps[k];
			bms_[k] = bms_seen_[k] ? 0.7 * bms_[k] + 0.3 * b : b;
			bms_seen_[k] = true;
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
		const bool work = p.n_cls > 0 || p.grid_blocks > 0 || p.n_fb > 0;
		if (busy)
			*busy = work || !pending_.empty();
		if (!work)
			return GRAYHIP_OK;
	}
	Result res;
	memset(res.res_claimed, 0, sizeof(res.res_claimed));
	memset(res.fresh_claimed, 0, sizeof(res.fresh_claimed));
	memset(res.saved, 0, sizeof(res.saved));
	res.fb_saved = res.g_res = res.g_fresh = res.g_saved = res.overflow = 0;
	res.steps[0] = res.steps[1] = res.steps[2] = 0;
	res.late[0] = res.late[1] = res.late[2] = 0;
	res.ms[0] = res.ms[1] = res.ms[2] = -1.0;
	int rc = dev_->epoch(p, res);
	std::lock_guard<std::mutex> lk(mu_);
	if (rc) {
		set_err(err, errlen, dev_->error());
		return rc;
	}
	/* every resumed ray and tile is claimed or carried over, and no list
	 * region overflowed */
	for (int k = 0; k < p.n_cls; ++k)
		if (res.res_claimed[k] < (unsigned)p.cls[k].n_resume) {
			set_err(err, errlen, "epoch left resumed rays behind");
			return GRAYHIP_EDEVICE;
		}
	if (p.grid_blocks > 0 && res.g_res < (unsigned)p.g_resume) {
		set_err(err, errlen, "epoch left resumed tiles behind");
		return GRAYHIP_EDEVICE;
	}
	if (res.overflow > 0) {
		set_err(err, errlen, "a resume or FP64 list region overflowed");
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
