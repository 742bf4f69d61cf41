/*
 * gray_grid.h -- the numerical-spacetime side of the gray job
 * (coordinates=dyst): HDF5 loading in the reference layout and the sliding
 * two-level time window of sim-org/gray.c:147-306, over the gridded device
 * path of include/grayhip.h.
 */
#ifndef GRAY_GRID_H
#define GRAY_GRID_H

#include <stddef.h>

#include "../../include/grayhip.h"

typedef struct gray_grid gray_grid;

int  gray_grid_create(const grayhip_config *cfg, gray_grid **out);
void gray_grid_destroy(gray_grid *g);
/* Open the file, list its time levels, read the axes and load the levels
 * bracketing t_init (or one level, frozen, when fast_light is set or the
 * file holds a single level). */
int  gray_grid_load(gray_grid *g, const char *path, double t_init, double dt_dump,
                    size_t i_init, size_t n_dump, size_t fast_light);
int  gray_grid_icond(gray_grid *g, double w_img, double h_img, double r_obs,
                     double i_obs, double j_obs);
/* One dump interval from t to target; slides the time window afterwards
 * when target leaves it (sim-org/gray.c:259-293). */
int  gray_grid_evolve(gray_grid *g, double t, double target, size_t n_sub, double *ns);
int  gray_grid_read(gray_grid *g, void *data, void *info);
const char *gray_grid_error(const gray_grid *g);

#endif
