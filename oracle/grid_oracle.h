/*
 * grid_oracle.h -- FP64 CPU restatement of the numerical-spacetime path.
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Grid arrays are interleaved per point: point index p = (k*ny + j)*nx + i
 * (x fastest, as the HDF5 datasets [z][y][x]); gam[l][p*40 + c] holds the
 * 40 Christoffels Gamma^mu_{ab} in file order (mu-major, then
 * tt tx ty tz xx xy xz yy yz zz), met[l][p*10 + c] the 10 metric
 * components g_tt g_tx g_ty g_tz g_xx g_xy g_xz g_yy g_yz g_zz, for the
 * two time levels l = 0 (t1) and 1 (t2).  Ray states are AoS with
 * og_grid_n_data reals per ray: q^mu, u^mu, then I_i and tau_i.
 */
#ifndef GRID_ORACLE_H
#define GRID_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "rt_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { OG_MAP_UNIFORM = 0, OG_MAP_FISHEYE = 1 };

/* Horizon proxy of the box-exit rule: stop where |dx/dt|^2 > 16. */
#define OG_GRID_U_CAP2 16.0

typedef struct {
	int    n[3];          /* nx, ny, nz                                   */
	double lo[3], hi[3];  /* first and last axis values                   */
	int    map;           /* OG_MAP_UNIFORM or OG_MAP_FISHEYE (n = 3)     */
	double t1, t2;        /* times of the two levels; equal: frozen       */
	double t_off;         /* grid time of the ray's q^t = 0               */
	int    raw_icond;     /* 1: no null normalisation (dyst.cl as written) */
	int    ref_rule;      /* 1: no box-exit stop (dyst.cl as written)      */
	const double *gam[2];
	const double *met[2];
	/* radiative transfer with the reference flow (flow.cl:36-58): density
	 * from the rho datasets, T_e = 1e12, fluid at rest, magnetic field
	 * b^mu = (0, 0, 0, b_field) (flow.cl:47 has b = 0: b_field = 0) */
	const double *rho[2];       /* per level, n_points; NULL without transfer */
	int    n_freq;              /* channels; state = 8 + 2 n_freq reals      */
	double nu[8];
	double m_adm;               /* M_ADM in solar masses (param.opts:4)      */
	const og_rt_consts *rt;
	double b_field;             /* B in Gauss; 0: the reference flow          */
} og_grid;

static inline int og_grid_n_data(const og_grid *G) { return 8 + 2 * G->n_freq; }

void    og_grid_down(const og_grid *G, const double q[4], const double u[4], double k[4]);
double  og_grid_getuu(const og_grid *G, const double q[4], const double u[4]);
void    og_grid_rhs(const og_grid *G, const double q[4], const double u[4], double du[4]);
void    og_grid_icond(const og_grid *G, double r_obs, double si, double ci, double sj,
                      double cj, double alpha, double beta, double q[4], double u[4]);
double  og_grid_getdt(const og_grid *G, const double q[4], const double u[4], double dt);
/* Full state derivative (geodesic and, with n_freq > 0, transfer). */
void    og_grid_rhs_all(const og_grid *G, const double *s, double *ds);
/* The flow at (q, u): density, and shift = -k.u_fluid with u_fluid at rest. */
void    og_grid_flow(const og_grid *G, const double q[4], const double u[4], double *ne,
                     double *shift, double *bkcos);
void    og_grid_integrate(const og_grid *G, double *s, double dt);
int64_t og_grid_evolve_ray(const og_grid *G, double *s, double dt, size_t n_sub, int *done);
void    og_grid_icond_all(const og_grid *G, size_t w_rays, size_t h_rays, double w_img,
                          double h_img, double r_obs, double si, double ci, double sj,
                          double cj, double *data, double *info);
int64_t og_grid_evolve_all(const og_grid *G, size_t n_rays, double *data, double *info,
                           double dt, size_t n_sub, int64_t *steps_out, int n_threads);
/* One interpolated field: which = 0 Christoffel comp (0..39), 1 metric
 * (0..9), 2 density. */
double  og_grid_sample(const og_grid *G, int which, int comp, const double q[4]);

#ifdef __cplusplus
}
#endif
#endif
