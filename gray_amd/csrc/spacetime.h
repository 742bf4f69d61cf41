/*
 * spacetime.h -- the resident numerical spacetime object behind
 * grayhip_spacetime (include/grayhip.h).
 */
#pragma once

#include "grid_device.h"

struct grayhip_spacetime {
	grayhip_grid_desc desc;
	size_t n_points;
	void *gam[2], *met[2];   /* slots; slot 0 = t1, slot 1 = t2             */
	void *rho[2];            /* density of each slot                       */
	double time[2];
	int loaded[2];
	int has_rho[2];
	int frozen;
	double *d_tmp;           /* scratch for sample() events / results       */
	size_t tmp_cap;
	char err[256];
};

/* Density present in every slot a render would read. */
int grayhip_spacetime_has_rho(const grayhip_spacetime *st);

/* Kernel-side description for a render with ray time offset t_off. */
gray::GridDesc spacetime_view(const grayhip_spacetime *st, double t_off, int raw_icond,
                              int ref_rule);
