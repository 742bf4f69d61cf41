/*
 * srv_sched.h -- host scheduler of the serving pool, independent of HIP.
 *
 * Holds every request the pool knows of and decides each epoch's work:
 * admission by the output budget (a first-fit arena) and by a free Kerr
 * class, the share of the Kerr and grid launches each class and kind gets,
 * and the queues of fresh rays and tiles, shared out round robin among
 * the requests still waiting to start, oldest first.  The GPU side is a
 * Device (the HIP back-end in grayhip_server.hip; tests drive a CPU mock
 * through the same interface).
 *
 * Threads: one thread calls step(); submit(), next_done(), state(),
 * result() and stats() may run on other threads at the same time (one
 * mutex guards the bookkeeping and is not held while an epoch runs).
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <deque>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/grayhip.h"
#include "srv_consts.h"

namespace gray {
namespace srv {

/* Parameters every request of a Kerr class shares (all but spin and
 * camera). */
struct ClassKey {
	double dt;
	size_t n_sub, n_dump;
	int n_freq, fluid, escape;
	double r_esc;
	double nu[SRV_NF];
	bool operator==(const ClassKey &o) const;
};

/* What the scheduler knows of a request (the device back-end also gets the
 * request as submitted). */
struct Req {
	int64_t id;
	int state;            /* GRAYHIP_REQ_*                                  */
	int tab;              /* device table index while admitted             */
	int kind;             /* SRV_KERR or SRV_GRID                          */
	int cls;              /* Kerr class while admitted                     */
	int nd;               /* 8 + 2 n_freq                                  */
	size_t n_rays;
	size_t units;         /* rays (Kerr) or 8x8 tiles (grid)               */
	size_t next;          /* first unit not yet handed to a queue          */
	size_t off, bytes;    /* output region in the arena                    */
	ClassKey key;
	grayhip_request q;
};

/* Fresh work of one request: units first .. first + count - 1. */
struct Seg {
	int tab;
	int first;
	unsigned count;
};

struct ClassPlan {
	ClassKey key;
	int slot;             /* scheduler's class index                        */
	int block0, block1;   /* blocks of the Kerr launch                      */
	int res_off, n_resume;/* resumed rays in the resume in-list             */
	int out_base, out_cap;/* the class's region of the resume out-list      */
	int seg0, seg1;       /* its segments in Plan::kseg                     */
	unsigned n_fresh;     /* fresh rays in them                             */
};

/* One epoch's work for the device.  The FP64 and grid launches start
 * first; the Kerr launch fills the room they leave.  Every list an epoch
 * reads is written anew into the other list of its pair (flip). */
struct Plan {
	int n_cls;
	ClassPlan cls[SRV_MAX_CLASSES];
	std::vector<Seg> kseg;
	int kerr_blocks, K_kerr;
	int grid_blocks, K_grid;
	int g_res_off, g_resume;   /* resumed tile records in the in-list       */
	int g_out_cap;             /* records of the out region                 */
	std::vector<Seg> gseg;
	unsigned g_fresh;          /* fresh tiles in gseg                        */
	int fb_blocks, K_fb;
	int fb_off, n_fb;          /* FP64 entries in the FP64 in-list           */
	double epoch_ms;           /* stop time of the launches (<= 0: their K)  */
	int flip;                  /* which list of each pair is the in-list     */
};

struct Result {
	unsigned res_claimed[SRV_MAX_CLASSES], fresh_claimed[SRV_MAX_CLASSES];
	unsigned saved[SRV_MAX_CLASSES];
	unsigned fb_saved, g_res, g_fresh, g_saved, overflow;
	std::vector<int> done;          /* table indices of finished requests    */
	unsigned long long steps[3];
	double ms[3];                   /* Kerr, FP64, grid kernel times (<0: none) */
	unsigned late[3];               /* waves that started after the stop time */
};

/* Device limits the scheduler plans against. */
struct Caps {
	int kerr_blocks;      /* Kerr blocks the GPU holds                        */
	int grid_blocks;      /* grid blocks the GPU holds (4 waves each)          */
	int n_cu;             /* compute units                                    */
	int kerr_beside;      /* Kerr blocks a CU holds beside one FP64 or grid
	                         block                                         */
	int res_cap;          /* entries of each Kerr resume list                 */
	int tile_cap;         /* records of each grid tile list                   */
	int fb_cap;           /* entries of each FP64 list                        */
	int max_segs;         /* fresh segments of one kind per epoch             */
	int max_requests;     /* device table size                               */
	size_t arena;         /* bytes of request outputs                        */
};

class Device {
public:
	virtual ~Device() {}
	/* Write table entry tab for r (outputs at arena offset r.off). */
	virtual int admit(int tab, const Req &r) = 0;
	virtual int epoch(const Plan &p, Result &res) = 0;
	/* Copy a finished request's outputs (any pointer may be NULL). */
	virtual int fetch(const Req &r, double *data, double *info, int64_t *steps) = 0;
	virtual const char *error() const = 0;
};

struct Config {
	int epoch_steps;      /* initial iterations per Kerr wave                */
	double epoch_ms;      /* target epoch duration for adapting the Ks; <= 0
	                         keeps them fixed                              */
};

class Sched {
public:
	Sched(const Config &cfg, const Caps &caps, Device *dev);
	/* r: kind, nd, n_rays, units, bytes, key and q filled in by the caller. */
	int submit(Req r, int64_t *id, char *err, size_t errlen);
	int step(int *busy, char *err, size_t errlen);
	int next_done(int64_t *id);
	int state(int64_t id, int *st);
	/* 0, GRAYHIP_EINVAL (unknown), GRAYHIP_ESTATE (not finished) or the
	 * device's code */
	int result(int64_t id, double *data, double *info, int64_t *steps, char *err, size_t errlen);
	void stats(int64_t steps[3], int64_t *epochs);

private:
	struct Arena {
		std::map<size_t, size_t> free_;   /* offset -> size */
		void init(size_t size);
		bool alloc(size_t bytes, size_t *off);
		void release(size_t off, size_t bytes);
	};
	struct Class {
		ClassKey key;
		int refs;                 /* admitted unfinished requests           */
		int n_resume, res_off;    /* rays saved at the last epoch (in the
		                             resume list that is next read)        */
	};

	int admit_locked(char *err, size_t errlen);
	void plan_locked(Plan &p);
	void consume_locked(const Plan &p, const Result &res);
	void fill_segs(std::vector<int64_t> &order, size_t want, size_t chunk, size_t max_segs,
	               std::vector<Seg> &out, unsigned *total);

	Config cfg_;
	Caps caps_;
	Device *dev_;
	std::mutex mu_;
	std::map<int64_t, Req> reqs_;
	std::deque<int64_t> pending_, done_;
	std::vector<int64_t> running_;        /* admitted, not finished (oldest first) */
	std::vector<int> free_tab_;
	std::vector<int64_t> tab_req_;        /* request id per table index (-1 free) */
	Arena arena_;
	Class cls_[SRV_MAX_CLASSES];
	bool cls_used_[SRV_MAX_CLASSES];
	int g_res_off_, g_resume_;            /* grid records saved last epoch     */
	int fb_off_, n_fb_;                   /* FP64 entries for the next epoch   */
	int flip_;
	int K_[3];                            /* Kerr, FP64, grid iterations       */
	/* work model of the two launches (0 Kerr, 1 grid) */
	double units_[2];                     /* units started                     */
	double bms_[2];                       /* block-ms per step                 */
	bool bms_seen_[2];
	double spu(int k) const;
	int64_t next_id_, epochs_;
	unsigned long long steps_[3];
};

} // namespace srv
} // namespace gray
