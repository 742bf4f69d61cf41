/*
 * gray_grid.c -- numerical spacetimes for the gray job (coordinates=dyst).
 */
#include "gray_grid.h"

#include <stdlib.h>

struct gray_grid {
	char err[256];
};

static const char *not_ready = "numerical spacetimes (coordinates=dyst) are not available in this build";

int gray_grid_create(const grayhip_config *cfg, gray_grid **out)
{
	(void)cfg;
	*out = NULL;
	return -1;
}
void gray_grid_destroy(gray_grid *g) { free(g); }
int gray_grid_load(gray_grid *g, const char *path, double t_init, double dt_dump, size_t i_init,
                   size_t n_dump, size_t fast_light)
{
	(void)g; (void)path; (void)t_init; (void)dt_dump; (void)i_init; (void)n_dump; (void)fast_light;
	return -1;
}
int gray_grid_icond(gray_grid *g, double w_img, double h_img, double r_obs, double i_obs, double j_obs)
{
	(void)g; (void)w_img; (void)h_img; (void)r_obs; (void)i_obs; (void)j_obs;
	return -1;
}
int gray_grid_evolve(gray_grid *g, double t, double target, size_t n_sub, double *ns)
{
	(void)g; (void)t; (void)target; (void)n_sub; (void)ns;
	return -1;
}
int gray_grid_read(gray_grid *g, void *data, void *info)
{
	(void)g; (void)data; (void)info;
	return -1;
}
const char *gray_grid_error(const gray_grid *g)
{
	(void)g;
	return not_ready;
}
