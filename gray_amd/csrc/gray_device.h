/*
 * gray_device.h -- per-ray device math of the GRay2 geodesic path for
 * gfx950, templated on the arithmetic type T (double or float).
 *
 * The double instantiation is compiled with -ffp-contract=off and evaluates
 * every expression in exactly the order written here, which is the order of
 * the reference kernels (sim-org/KS.cl, phys.cl, RK4.cl, driver.cl) with
 * OpenCL dot() taken left to right.  Division and sqrt in double are the
 * correctly rounded IEEE operations, so the plain FP64 form reproduces a
 * correctly rounded CPU evaluation of the same sequence bit for bit.
 *
 * The geodesic RHS (evaluated four times per step) fuses its multiply-adds
 * explicitly, as an OpenCL compiler may under the default FP_CONTRACT ON
 * (KS.cl counts them, "35 (-9) FLOPs"), by one fixed rule: in a sum taken
 * left to right, a product added to the running value is fused into it
 * (fmaT); where both operands of an addition are products, the left one is
 * rounded and the right one fused.  The once-per-ray camera set-up, down(),
 * getuu() and the step rule are evaluated unfused.
 *
 * The float instantiation is compiled with contraction on and uses the
 * hardware reciprocal / square-root / exp2 instructions (about 1 ulp).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "fpdiv.h"
#include "grayhip_params.h"

namespace gray {

template <class T> struct Mth;

template <> struct Mth<double> {
	static __device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
	static __device__ __forceinline__ double div(double a, double b) { return a / b; }
	static __device__ __forceinline__ double exp_(double x) { return exp(x); }
};

template <> struct Mth<float> {
	static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
	static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
	static __device__ __forceinline__ float exp_(float x) { return __expf(x); }
};

template <class T> struct V4 { T t, x, y, z; };

template <class T>
__device__ __forceinline__ T dot4(const V4<T> &a, const V4<T> &b)
{
	return ((a.t * b.t + a.x * b.x) + a.y * b.y) + a.z * b.z;
}

/* a b + c with one rounding (v_fma_f64 / v_fma_f32). */
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

/* dot4 with every product after the first fused into the running sum. */
template <class T>
__device__ __forceinline__ T dot4f(const V4<T> &a, const V4<T> &b)
{
	return fmaT(a.z, b.z, fmaT(a.y, b.y, fmaT(a.x, b.x, a.t * b.t)));
}

/* sim-org/KS.cl:48-55 */
template <class T>
__device__ __forceinline__ T getrr(T a, const V4<T> &q)
{
	T aa = a * a;
	T zz = q.z * q.z;
	T kk = T(0.5) * (((q.x * q.x + q.y * q.y) + zz) - aa);
	return Mth<T>::sqrt_(kk * kk + aa * zz) + kk;
}

/* sim-org/KS.cl:91-114: k_mu = g_{mu nu} u^nu, g = eta + f l l. */
template <class T>
__device__ __forceinline__ V4<T> down(T a, const V4<T> &q, const V4<T> &u)
{
	typedef Mth<T> M;
	T aa = a * a;
	T zz = q.z * q.z;
	T kk = T(0.5) * (((q.x * q.x + q.y * q.y) + zz) - aa);
	T rr = M::sqrt_(kk * kk + aa * zz) + kk;
	T r  = M::sqrt_(rr);

	T f  = M::div((T(2.0) * rr) * r, rr * rr + aa * zz);
	T lx = M::div(r * q.x + a * q.y, rr + aa);
	T ly = M::div(r * q.y - a * q.x, rr + aa);
	T lz = M::div(q.z, r);

	V4<T> gt = {T(-1.0) + f, f * lx, f * ly, f * lz};
	V4<T> gx = {f * lx, T(1.0) + (f * lx) * lx, (f * lx) * ly, (f * lx) * lz};
	V4<T> gy = {f * ly, (f * ly) * lx, T(1.0) + (f * ly) * ly, (f * ly) * lz};
	V4<T> gz = {f * lz, (f * lz) * lx, (f * lz) * ly, T(1.0) + (f * lz) * lz};
	V4<T> k = {dot4(gt, u), dot4(gx, u), dot4(gy, u), dot4(gz, u)};
	return k;
}

template <class T>
__device__ __forceinline__ T getuu(T a, const V4<T> &q, const V4<T> &u)
{
	return dot4(down(a, q, u), u);
}

/* The inverse of down(): u^mu = g^{mu nu} p_nu with g^{mu nu} = eta^{mu nu}
 * - f l^mu l^nu, l^mu = (-1, lx, ly, lz) (l is null), so u^mu = eta^{mu nu}
 * p_nu - f (l^nu p_nu) l^mu; f and l as in down(). */
template <class T>
__device__ __forceinline__ V4<T> up(T a, const V4<T> &q, const V4<T> &p)
{
	typedef Mth<T> M;
	T aa = a * a;
	T zz = q.z * q.z;
	T kk = T(0.5) * (((q.x * q.x + q.y * q.y) + zz) - aa);
	T rr = M::sqrt_(kk * kk + aa * zz) + kk;
	T r  = M::sqrt_(rr);

	T f  = M::div((T(2.0) * rr) * r, rr * rr + aa * zz);
	T lx = M::div(r * q.x + a * q.y, rr + aa);
	T ly = M::div(r * q.y - a * q.x, rr + aa);
	T lz = M::div(q.z, r);

	T fL = f * (((lx * p.x + ly * p.y) + lz * p.z) - p.t);
	V4<T> u = {fL - p.t, p.x - fL * lx, p.y - fL * ly, p.z - fL * lz};
	return u;
}

/* sim-org/KS.cl:188-232 with the camera trigonometry passed in. */
template <class T>
__device__ __forceinline__ void gr_icond(T a, T r_obs, T si, T ci, T sj, T cj,
                                         T alpha, T beta, V4<T> &q, V4<T> &u)
{
	typedef Mth<T> M;
	T R0 = r_obs * si - beta * ci;
	T z  = r_obs * ci + beta * si;
	T y  = R0 * sj + alpha * cj;
	T x  = R0 * cj - alpha * sj;
	q = V4<T>{T(0.0), x, y, z};
	u = V4<T>{T(1.0), si * cj, si * sj, ci};

	T aa = a * a;
	T zz = z * z;
	T kk = T(0.5) * (((x * x + y * y) + zz) - aa);
	T rr = M::sqrt_(kk * kk + aa * zz) + kk;
	T r  = M::sqrt_(rr);
	T f  = M::div((T(2.0) * rr) * r, rr * rr + aa * zz);
	T lx = M::div(r * x + a * y, rr + aa);
	T ly = M::div(r * y - a * x, rr + aa);
	T lz = M::div(z, r);

	T g01 = f * lx, g02 = f * ly, g03 = f * lz;
	T g11 = T(1.0) + (f * lx) * lx, g12 = (f * lx) * ly, g13 = (f * lx) * lz;
	T g21 = (f * ly) * lx, g22 = T(1.0) + (f * ly) * ly, g23 = (f * ly) * lz;
	T g31 = (f * lz) * lx, g32 = (f * lz) * ly, g33 = T(1.0) + (f * lz) * lz;

	T A = T(-1.0) + f;
	T B = ((g01 * u.x + g02 * u.y) + g03 * u.z) * T(2.0);
	T C = ((((g11 * u.x + g12 * u.y) + g13 * u.z) * u.x) +
	       (((g21 * u.x + g22 * u.y) + g23 * u.z) * u.y)) +
	       (((g31 * u.x + g32 * u.y) + g33 * u.z) * u.z);
	T s = M::div(-(B + M::sqrt_(B * B - (T(4.0) * A) * C)), T(2.0) * A);
	u.x = M::div(u.x, s);
	u.y = M::div(u.y, s);
	u.z = M::div(u.z, s);
}

/* Kerr-Schild quantities at the event that gr_rhs computes anyway: r, f and
 * the covariant l_mu = (1, lx, ly, lz), and l.u.  With them the metric is
 * g = eta + f l l, so k_mu = eta_mu nu u^nu + f (l.u) l_mu needs no second
 * evaluation (used by the FP32 form's transfer). */
template <class T> struct KsAux {
	T r, ir, f, lx, ly, lz, lu;   /* ir = 1 / r */
	T R2, zz, B;                  /* x^2 + y^2, z^2, y p_x - x p_y: ks_ham only */
};

/* sim-org/KS.cl:263-378: regrouped Cartesian Kerr-Schild geodesic RHS in
 * FP64 (the FP32 form integrates the Hamiltonian form below).  Returns
 * du/dlambda (dq/dlambda = u); fills *aux when given.  M supplies sqrt and
 * division (the correctly rounded ones by default). */
template <class T, class M = Mth<T>>
__device__ __forceinline__ V4<T> gr_rhs(T a, const V4<T> &q, const V4<T> &u,
                                        KsAux<T> *aux = nullptr)
{
	static_assert(sizeof(T) == 8, "gr_rhs is the FP64 RHS");
	T f, dx_f, dy_f, dz_f;
	T lx, dx_lx, dy_lx, dz_lx;
	T ly, dx_ly, dy_ly, dz_ly;
	T lz, dx_lz, dy_lz, dz_lz;
	T tmp;
	{
		T aa = a * a;
		T zz = q.z * q.z;
		T kk = T(0.5) * ((fmaT(q.y, q.y, q.x * q.x) + zz) - aa);
		T dd, rr, r, ir, half_rdd, two_rq, iss;
		if constexpr (std::is_same<M, Mth<double>>::value) {
			/* The correctly rounded operations as sqrt_core / div_core,
			 * which give the same bits where every operand lies in the
			 * ranges checked below (d2 and rr in [2^-600, 2^600] put r,
			 * dd, r dd and rr + aa z^2/rr in [2^-600, 2^701]; a^2 z^2 = 0
			 * or in [2^-400, 2^100] keeps a^2 z^2 / rr normal); any other
			 * lane evaluates them again with the full sequences. */
			const T d2 = fmaT(aa, zz, kk * kk);
			const T azz = aa * zz;
			dd = sqrt_core(d2);
			rr = dd + kk;
			r  = sqrt_core(rr);
			ir = div_core(T(1.0), r);
			half_rdd = div_core(T(0.5), r * dd);
			two_rq = div_core(T(2.0), rr + div_core(azz, rr));
			iss = div_core(T(1.0), rr + aa);
			if (!(in_range(d2, 0x1p-600, 0x1p600) && in_range(rr, 0x1p-600, 0x1p600) &&
			      (azz == 0.0 || in_range(azz, 0x1p-400, 0x1p100)) && aa <= 0x1p100)) {
				dd = M::sqrt_(d2);
				rr = dd + kk;
				r  = M::sqrt_(rr);
				ir = M::div(T(1.0), r);
				half_rdd = M::div(T(0.5), r * dd);
				two_rq = M::div(T(2.0), rr + M::div(azz, rr));
				iss = M::div(T(1.0), rr + aa);
			}
		} else {
			dd = M::sqrt_(fmaT(aa, zz, kk * kk));
			rr = dd + kk;
			r  = M::sqrt_(rr);
			ir = M::div(T(1.0), r);
			half_rdd = M::div(T(0.5), r * dd);
			two_rq = M::div(T(2.0), rr + M::div(aa * zz, rr));
			iss = M::div(T(1.0), rr + aa);
		}
		if (aux) {
			aux->r = r;
			aux->ir = ir;
		}
		T ss = rr + aa;
		T dx_r, dy_r, dz_r, tmp2;

		tmp  = half_rdd;
		dz_r = (tmp * ss) * q.z;
		tmp  = tmp * rr;
		dy_r = tmp * q.y;
		dx_r = tmp * q.x;
		tmp  = two_rq;

		tmp2 = fmaT(-(T(2.0) * rr), tmp, T(3.0));
		f    = tmp * r;
		dx_f = (tmp * dx_r) * tmp2;
		dy_f = (tmp * dy_r) * tmp2;
		dz_f = tmp * fmaT(-((tmp * aa) * q.z), ir, dz_r * tmp2);

		T m2r  = T(-2.0) * r;
		T issr = iss * r;
		T issa = iss * a;

		lx    = iss * fmaT(q.y, a, q.x * r);
		tmp   = iss * fmaT(m2r, lx, q.x);
		dx_lx = fmaT(tmp, dx_r, issr);
		dy_lx = fmaT(tmp, dy_r, issa);
		dz_lx = tmp * dz_r;

		ly    = iss * fmaT(-q.x, a, q.y * r);
		tmp   = iss * fmaT(m2r, ly, q.y);
		dx_ly = fmaT(tmp, dx_r, -issa);
		dy_ly = fmaT(tmp, dy_r, issr);
		dz_ly = tmp * dz_r;

		lz    = q.z * ir;
		tmp   = (-lz) * ir;
		dx_lz = tmp * dx_r;
		dy_lz = tmp * dy_r;
		dz_lz = fmaT(tmp, dz_r, ir);
	}

	T hDxu, hDyu, hDzu;
	V4<T> uD;
	{
		V4<T> Dx, Dy, Dz;
		T lu  = fmaT(lz, u.z, fmaT(ly, u.y, fmaT(lx, u.x, u.t)));
		if (aux) {
			aux->f = f;
			aux->lx = lx;
			aux->ly = ly;
			aux->lz = lz;
			aux->lu = lu;
		}
		T flu = f * lu;
		Dx.t = fmaT(f, fmaT(dx_lz, u.z, fmaT(dx_ly, u.y, dx_lx * u.x)), dx_f * lu);
		Dy.t = fmaT(f, fmaT(dy_lz, u.z, fmaT(dy_ly, u.y, dy_lx * u.x)), dy_f * lu);
		Dz.t = fmaT(f, fmaT(dz_lz, u.z, fmaT(dz_ly, u.y, dz_lx * u.x)), dz_f * lu);

		Dx.x = fmaT(flu, dx_lx, Dx.t * lx);
		Dx.y = fmaT(flu, dx_ly, Dx.t * ly);
		Dx.z = fmaT(flu, dx_lz, Dx.t * lz);

		Dy.x = fmaT(flu, dy_lx, Dy.t * lx);
		Dy.y = fmaT(flu, dy_ly, Dy.t * ly);
		Dy.z = fmaT(flu, dy_lz, Dy.t * lz);

		Dz.x = fmaT(flu, dz_lx, Dz.t * lx);
		Dz.y = fmaT(flu, dz_ly, Dz.t * ly);
		Dz.z = fmaT(flu, dz_lz, Dz.t * lz);

		hDxu = T(0.5) * dot4f(Dx, u);
		hDyu = T(0.5) * dot4f(Dy, u);
		hDzu = T(0.5) * dot4f(Dz, u);

		uD.t = fmaT(u.z, Dz.t, fmaT(u.y, Dy.t, u.x * Dx.t));
		uD.x = fmaT(u.z, Dz.x, fmaT(u.y, Dy.x, u.x * Dx.x));
		uD.y = fmaT(u.z, Dz.y, fmaT(u.y, Dy.y, u.x * Dx.y));
		uD.z = fmaT(u.z, Dz.z, fmaT(u.y, Dy.z, u.x * Dx.z));

		tmp = f * fmaT(lz, uD.z - hDzu, fmaT(ly, uD.y - hDyu, fmaT(lx, uD.x - hDxu, -uD.t)));
	}
	V4<T> acc = {uD.t - tmp, fmaT(lx, tmp, hDxu - uD.x),
	             fmaT(ly, tmp, hDyu - uD.y), fmaT(lz, tmp, hDzu - uD.z)};
	return acc;
}

/* The FP32 form's geodesic RHS (and its FP64 re-integration's): the
 * Hamiltonian form of the same Kerr-Schild geodesics (sim-org/KS.cl:263-378
 * integrates u^mu), with the covariant momentum p_mu = g_{mu nu} u^nu as
 * the state.  H = g^{mu nu} p_mu p_nu / 2,
 * g^{mu nu} = eta^{mu nu} - f l^mu l^nu, l^mu = (-1, lx, ly, lz), L = l^mu p_mu:
 *   dx^mu/dlambda = dH/dp_mu = eta^{mu nu} p_nu - f L l^mu,
 *   dp_i/dlambda = -dH/dx^i = L (L d_i f / 2 + f d_i(l^nu p_nu)),
 *   dp_t/dlambda = 0 (the metric is stationary),
 * which needs f, l and first derivatives but no Christoffel contraction.
 * With lx = (r x + a y) / (r^2 + a^2), ly = (r y - a x) / (r^2 + a^2),
 * lz = z / r, A = x p_x + y p_y, B = y p_x - x p_y, S = lx p_x + ly p_y:
 *   d_i(l^nu p_nu) = d_i r K1 + H_i,  K1 = (A - 2 r S) / (r^2 + a^2) - lz p_z / r,
 *   H = ((r p_x - a p_y), (r p_y + a p_x)) / (r^2 + a^2), p_z / r;
 *   d_i f = F1 d_i r (+ F2 for i = z), d_i r = c x_i (c = r / (2 r dd)
 *   times r^2 for x, y and r^2 + a^2 for z),
 * so dp_i = L (x_i c M + f H_i) (+ L^2 F2 / 2 for z), M = L F1 / 2 + f K1.
 * r, 1/r and 1/(r dd) come from two reciprocal square roots and f and
 * 1/(r^2 + a^2) from two reciprocals, R's rsq and rcp (HwF32: the hardware
 * instructions, ~1 ulp; FP64: refined by Newton steps); 2 / (rr + aa zz /
 * rr) = 2 rr / (rr^2 + aa zz).  Fills *aux (aux->lu = L = l_mu u^mu, as l is
 * null). */
struct HwF32 {
	static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
	static __device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
};

template <class T> struct Ham {
	V4<T> u;                 /* dx^mu / dlambda     */
	T dpx, dpy, dpz;         /* dp_i / dlambda      */
	T r;                     /* Kerr-Schild r of q  */
};

template <class T, class R>
__device__ __forceinline__ Ham<T> ks_ham(T a, const V4<T> &q, const V4<T> &p,
                                         KsAux<T> *aux = nullptr)
{
	const T aa = a * a;
	const T zz = q.z * q.z;
	const T R2 = fmaT(q.y, q.y, q.x * q.x);
	const T kk = T(0.5) * ((R2 + zz) - aa);
	const T d2 = fmaT(aa, zz, kk * kk);
	const T idd = R::rsq(d2);
	const T dd = d2 * idd;
	const T rr = dd + kk;
	const T ir = R::rsq(rr);
	const T r = rr * ir;
	const T half_rdd = (T(0.5) * ir) * idd;
	const T two_rq = (T(2.0) * rr) * R::rcp(fmaT(aa, zz, rr * rr));
	const T ss = rr + aa;
	const T iss = R::rcp(ss);

	const T f = two_rq * r;
	const T F1 = two_rq * fmaT(-(T(2.0) * rr), two_rq, T(3.0));
	const T F2 = -two_rq * (((two_rq * aa) * q.z) * ir);
	const T lx = iss * fmaT(q.y, a, q.x * r);
	const T ly = iss * fmaT(-q.x, a, q.y * r);
	const T lz = q.z * ir;

	const T A = fmaT(q.y, p.y, q.x * p.x);
	const T B = fmaT(q.y, p.x, -(q.x * p.y));
	const T S = iss * fmaT(r, A, a * B);
	const T lzpz = lz * p.z;
	const T L = (S + lzpz) - p.t;
	const T fL = f * L;
	Ham<T> h;
	h.u = V4<T>{fL - p.t, fmaT(-fL, lx, p.x), fmaT(-fL, ly, p.y), fmaT(-fL, lz, p.z)};

	const T K1 = fmaT(-lzpz, ir, iss * fmaT(T(-2.0) * r, S, A));
	const T hL = T(0.5) * L;
	const T Mc = fmaT(hL, F1, f * K1);
	const T Mxy = (half_rdd * rr) * Mc;
	const T Mz = (half_rdd * ss) * Mc;
	const T hx = iss * fmaT(r, p.x, -(a * p.y));
	const T hy = iss * fmaT(r, p.y, a * p.x);
	const T hz = p.z * ir;
	h.dpx = L * fmaT(q.x, Mxy, f * hx);
	h.dpy = L * fmaT(q.y, Mxy, f * hy);
	h.dpz = L * fmaT(hL, F2, fmaT(q.z, Mz, f * hz));
	h.r = r;
	if (aux) {
		aux->r = r;
		aux->ir = ir;
		aux->f = f;
		aux->lx = lx;
		aux->ly = ly;
		aux->lz = lz;
		aux->lu = L;
		aux->R2 = R2;
		aux->zz = zz;
		aux->B = B;
	}
	return h;
}

/* sim-org/phys.cl:48-64 with the selectable escape rule, for the
 * Kerr-Schild r of q given.  P is any parameter block with the fields a,
 * r_h, eps_stop, escape, r_esc (the render's RayParams, or per-lane
 * parameters of a serving pool). */
template <class T, class P>
__device__ __forceinline__ T getdt_r(const P &p, T r, const V4<T> &q, const V4<T> &u, T dt)
{
	T eps = r - T(p.r_h);

	if (eps < T(p.eps_stop))
		return T(0.0);
	if (p.escape == GRAYHIP_ESC_OUTWARD) {
		T qu = (q.x * u.x + q.y * u.y) + q.z * u.z;
		if (r > T(p.r_esc) && qu * dt > T(0.0))
			return T(0.0);
	} else {
		T qq = (q.x * q.x + q.y * q.y) + q.z * q.z;
		if (qq < T(0.0) && r > T(1e3))
			return T(0.0);
	}
	T cap = (T(0.01) * eps) * eps;
	if (fabs(dt) > cap)
		return dt > T(0.0) ? cap : -cap;
	return dt;
}

/* getdt_r with r from q (sim-org/phys.cl:48-64). */
template <class T, class P>
__device__ __forceinline__ T getdt(const P &p, const V4<T> &q,
                                   const V4<T> &u, T dt)
{
	T r;
	if constexpr (sizeof(T) == 8) {
		/* sqrt_core gives the correctly rounded root for d2 and rr in
		 * [2^-600, 2^600]; other lanes take the full sequence. */
		const T a = p.a, aa = a * a, zz = q.z * q.z;
		const T kk = T(0.5) * (((q.x * q.x + q.y * q.y) + zz) - aa);
		const T d2 = kk * kk + aa * zz;
		const T rr = sqrt_core(d2) + kk;
		r = sqrt_core(rr);
		if (!(in_range(d2, 0x1p-600, 0x1p600) && in_range(rr, 0x1p-600, 0x1p600)))
			r = Mth<T>::sqrt_(getrr(a, q));
	} else {
		r = Mth<T>::sqrt_(getrr(T(p.a), q));
	}
	return getdt_r(p, r, q, u, dt);
}

} // namespace gray
