/*
 * gray_oracle.h -- FP64 CPU restatement of the GRay2 geodesic hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke entry point and the CPU-baseline leg of the benchmark may
 * load it.  The product path (libgrayhip.so) never links or calls it.
 *
 * Every function restates one piece of the reference (luxsrc/gray, sim-org/)
 * from its equations, with the evaluation order written out explicitly so
 * that the FP64 GPU kernels can reproduce it bit for bit:
 *   - plain IEEE-754 double operations, no fused multiply-add
 *     (built with -ffp-contract=off);
 *   - sqrt and division correctly rounded (the only non-trivial ops on the
 *     geodesic path);
 *   - OpenCL dot(a,b) is evaluated left to right: ((a0*b0+a1*b1)+a2*b2)+a3*b3;
 *   - the camera trigonometry sin/cos(i_obs), sin/cos(j_obs) is computed
 *     once on the host and passed in.
 *
 * State layout per ray (n_data = 8 + 2*n_freq reals, sim-org/phys.cl:30-33,
 * rt.cl:174-177):  t x y z  u^t u^x u^y u^z  I[0..n_freq)  tau[0..n_freq)
 */
#ifndef GRAY_ORACLE_H
#define GRAY_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OG_MAX_FREQ 8

/* Termination rules.  OG_ESC_REF keeps sim-org/phys.cl:57 as written (the
 * escape test there can never fire); OG_ESC_OUTWARD stops a ray once its
 * Kerr-Schild radius exceeds r_esc while it moves outward along the
 * integration direction. */
enum { OG_ESC_REF = 0, OG_ESC_OUTWARD = 1 };

/* Fluid models feeding rt_rhs.  OG_FLUID_UNIFORM (n = 1, shift = 1
 * everywhere) exists to pin the transfer integrator against the slab
 * closed form I = (j/alpha)(1 - exp(-alpha L)). */
enum { OG_FLUID_NONE = 0, OG_FLUID_DISK = 1, OG_FLUID_UNIFORM = 2 };

typedef struct {
	double a;          /* dimensionless spin a_spin (M = 1)                */
	double r_h;        /* horizon radius 1 + sqrt(1 - a^2)                   */
	double eps_stop;   /* stop when r - r_h < eps_stop (ref: 0.01)           */
	int    escape;     /* OG_ESC_REF or OG_ESC_OUTWARD                       */
	double r_esc;      /* escape radius for OG_ESC_OUTWARD                   */
	int    n_freq;     /* number of frequency channels (0: RT off)           */
	double nu[OG_MAX_FREQ];
	int    fluid;      /* OG_FLUID_*                                         */
	double r_isco;     /* inner edge of the synthetic disk                   */
} og_params;

/* Fill r_h from a (1 + sqrt(1 - a*a), sim-org/KS.cl:60). */
void   og_setup(og_params *p);

/* Prograde ISCO radius (Bardeen, Press & Teukolsky 1972), M = 1. */
double og_r_isco(double a);

/* sim-org/KS.cl:48-55 */
double og_getrr(double a, const double q[4]);
/* sim-org/KS.cl:57-61 */
double og_geteps(double a, const double q[4]);
/* sim-org/KS.cl:91-114: k_mu = g_{mu nu} u^nu */
void   og_down(double a, const double q[4], const double u[4], double k[4]);
/* sim-org/KS.cl:164-168 */
double og_getuu(double a, const double q[4], const double u[4]);
/* sim-org/KS.cl:188-232 (trigonometry passed in) */
void   og_gr_icond(double a, double r_obs, double si, double ci,
                   double sj, double cj, double alpha, double beta,
                   double q[4], double u[4]);
/* sim-org/KS.cl:263-378: (dq, du) = (u, du/dlambda) */
void   og_gr_rhs(double a, const double q[4], const double u[4],
                 double dq[4], double du[4]);
/* sim-org/phys.cl:48-64 with the termination rule in p */
double og_getdt(const og_params *p, const double q[4], const double u[4],
                double dt);
/* Synthetic disk flow: density n and redshift factor shift = -k.u_fluid. */
void   og_flow_disk(const og_params *p, const double q[4], const double k[4],
                    double *n, double *shift);
/* sim-org/phys.cl:66-74 + rt.cl:185-198: full state derivative */
void   og_rhs(const og_params *p, const double *s, double *ds);
/* sim-org/RK4.cl:37-47 */
void   og_integrate(const og_params *p, double *s, double dt);

/* sim-org/driver.cl:36-67 for one pixel (gi column, gj row). */
void   og_icond_ray(const og_params *p, size_t w_rays, size_t h_rays,
                    double w_img, double h_img, double r_obs,
                    double si, double ci, double sj, double cj,
                    size_t gi, size_t gj, double *state);
/* sim-org/driver.cl:70-108 for one ray and one launch; returns the number
 * of RK4 steps taken.  *done is set when the step rule returned zero. */
int64_t og_evolve_ray(const og_params *p, double *state, double dt,
                      size_t n_sub, int *done);

/* Whole-image drivers over AoS state [g][n_data], g = gi + gj*w_rays.
 * OpenMP over rays.  Returns total steps; steps_out (may be NULL) receives
 * per-ray step counts (accumulated), info receives g(u,u) per ray. */
void    og_icond_all(const og_params *p, size_t w_rays, size_t h_rays,
                     double w_img, double h_img, double r_obs,
                     double si, double ci, double sj, double cj,
                     double *data, double *info);
int64_t og_evolve_all(const og_params *p, size_t n_rays, double *data,
                      double *info, double dt, size_t n_sub,
                      int64_t *steps_out, int n_threads);
/* Boyer-Lindquist view of a Kerr-Schild state (KS.cl:63-149): spherical
 * Kerr-Schild position (t, r, theta, phi) and its BL form, with t matched
 * at r = r_match and phi at r = infinity; the velocity in spherical KS and
 * in BL.  Spheroidal convention x + i y = (r + i a) e^{i phi} sin(theta),
 * the one in which the metric of KS.cl:100-108 maps to BL by the r-only
 * shifts of KS.cl:79-88 (KS.cl:68-69,125 use the mirror image, r - i a).
 * KS.cl:146,149 multiply u^phi and u^t where the r-derivative of the
 * position map requires u^r; this follows the derivative:
 * u^t_BL = u^t - 2 r u^r / Delta, u^phi_BL = u^phi - a u^r / Delta. */
void   og_getsphks(double a, const double q[4], double out[4]);
void   og_getbl(double a, double r_match, const double q[4], double out[4]);
void   og_getsphksu(double a, const double q[4], const double u[4], double out[4]);
void   og_getblu(double a, const double q[4], const double u[4], double out[4]);

/* Evolve a strided subset of rays (every stride-th ray starting at 0); used
 * for bounded CPU-baseline timing.  Returns total steps. */
int64_t og_evolve_subset(const og_params *p, size_t n_rays, size_t stride,
                         double *data, double dt, size_t n_sub,
                         size_t n_dump, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
