/*
 * rt_oracle.h -- FP64 CPU restatement of the reference's radiative-transfer
 * physics (sim-org/rt.cl:28-198) and flow model (sim-org/flow.cl:36-58).
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Units are those of rt.cl: cgs, temperatures as theta_e = kT / (m_e c^2),
 * lengths of the transfer in units of G M_ADM M_sun / c^2 (hence the M_ADM
 * factor in the emissivities).
 *
 * Every constant that involves a transcendental function is computed once
 * (og_rt_consts) and handed to the device the same way, so the oracle and
 * the GPU evaluate the same expressions from the same constants.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OG_RT_EPS     1e-28   /* preamble.cl:30 */
#define OG_RT_T_MIN   1e-1    /* rt.cl:41-43: log K2 table range and size */
#define OG_RT_T_MAX   1e+2
#define OG_RT_T_GRID  60
#define OG_RT_TE_FLOW 1e12    /* flow.cl:51 */

typedef struct {
	double planck1;   /* 2 h c                                            */
	double planck2;   /* h / (m_e c)                                      */
	double c;         /* speed of light                                   */
	double gaunt_a;   /* sqrt(3 / pi)                                     */
	double gaunt_b;   /* sqrt(3) / pi                                     */
	double gaunt_c;   /* ln(4 / zeta), zeta = e^gamma_Euler = 1.78107...  */
	double gaunt_d;   /* ln(4 / zeta^2.5)                                 */
	double gaunt_e;   /* sqrt(12)                                         */
	double ff_x;      /* m_e c^2 / Ry                                     */
	double ff_y;      /* h / (m_e c^2)                                    */
	double ff_f;      /* sqrt(G M_sun / c^2 * 6.8e-38 / (4 pi sqrt(m_e c^2 / k_B))) */
	double syn_nus;   /* e / (9 pi m_e c)                                 */
	double syn_f;     /* (G M_sun / c^2) sqrt(2) pi e^2 / (3 c)           */
	double syn_k;     /* 2^(11/12)                                        */
	double k2_h;      /* T_GRID / ln(T_MAX / T_MIN)                       */
	double log_k2[OG_RT_T_GRID + 1]; /* ln K_2(1/theta_i), theta_i = T_MIN (T_MAX/T_MIN)^(i/T_GRID) */
} og_rt_consts;

/* ln K_2(x), modified Bessel function of the second kind, by the
 * trapezoidal rule on K_2(x) = int_0^inf exp(-x cosh t) cosh(2t) dt. */
double og_log_k2(double x);
void   og_rt_consts_init(og_rt_consts *c);

double og_log_k2it(const og_rt_consts *c, double te);
double og_b_planck(const og_rt_consts *c, double nu, double te);
double og_gaunt(const og_rt_consts *c, double x, double y);
double og_j_ff(const og_rt_consts *c, double nu, double te, double ne, double m_adm);
double og_j_syn(const og_rt_consts *c, double nu, double te, double ne, double b,
                double cos_theta, double m_adm);

/* rt.cl:185-198: dI_i/dlambda and dtau_i/dlambda for n_freq channels. */
void   og_rt_thermal(const og_rt_consts *c, int n_freq, const double *nu, double m_adm,
                     double ne, double te, double b, double bkcos, double shift,
                     const double *I, const double *tau, double *dI, double *dtau);

#ifdef __cplusplus
}
#endif
#endif
