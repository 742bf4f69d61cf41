/*
 * kernels_f64.hip -- FP64 kernels: camera initial conditions and the plain
 * FP64 evolve (one lane per ray, any number of intervals per launch).
 *
 * Built with -ffp-contract=off: the double arithmetic follows the written
 * order exactly (see gray_device.h).
 */
#include <climits>

#include "launch.h"
#include "step_f64.h"

namespace gray {

template <int NF>
__global__ __launch_bounds__(256) void k_icond(RayParams p, double *__restrict__ data,
                                               double *__restrict__ info,
                                               int64_t *__restrict__ steps, size_t n_rays)
{
	constexpr int ND = 8 + 2 * NF;
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	double s[ND];
	icond_ray<NF>(p, g, s);
#pragma unroll
	for (int i = 0; i < ND; ++i)
		data[(size_t)i * n_rays + g] = s[i];
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	info[g] = getuu(p.a, q, u);
	steps[g] = 0;
}

/* Plain FP64 form: one lane per ray, sr.n_dump intervals per launch in one
 * flat loop.  Each iteration takes one RK4 step of the lane's own
 * (interval, unit) position with interval_f64's quantisation, so a ray on a
 * long dependency chain near the photon orbit holds neither the other lanes
 * of its wave nor the later intervals of other rays (a launch per interval
 * would wait for it at every interval's end).  The results are those of
 * sr.n_dump one-interval launches bit for bit: at each interval boundary
 * the channels' shared depth takes the round trip of the stored form
 * (tau_0 = d_0 T stored, T = tau_0 / d_0 loaded), and a retired ray, which
 * a later launch would retire again at once, takes the remaining round
 * trips without steps. */
template <int NF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NF <= 4 ? 2 : 1))) void k_evolve_f64(RayParams p, StepRule sr,
                                                    double *__restrict__ data,
                                                    double *__restrict__ info,
                                                    int64_t *__restrict__ steps,
                                                    size_t n_rays)
{
	constexpr int ND = nd64<NF>();
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	const RtConst64<NF> rc = make_rtconst64<NF>(p);
	/* stored: 8 + 2 NF reals with tau_i in place; integrated: T = tau_0 / d_0 */
	double s[ND];
#pragma unroll
	for (int i = 0; i < ND; ++i)
		s[i] = data[(size_t)i * n_rays + g];
	if constexpr (NF > 0)
		s[8 + NF] = s[8 + NF] / rc.d[0];
	int64_t st = 0;
	int d = 0, h = 0;
	while (d < sr.n_dump) {
		V4<double> q = {s[0], s[1], s[2], s[3]};
		V4<double> u = {s[4], s[5], s[6], s[7]};
		double x = getdt(p, q, u, sr.dts) / sr.dt * (double)sr.n;
		int dh = (x == x) ? (int)x : 0;
		if (!dh) {
			/* retired: every later interval would retire at its first
			 * step rule again */
			if constexpr (NF > 0) {
				for (++d; d < sr.n_dump; ++d)
					s[8 + NF] = (rc.d[0] * s[8 + NF]) / rc.d[0];
			}
			break;
		}
		if (dh > sr.n - h)
			dh = sr.n - h;
		rk4_f64<NF>(p, rc, s, (double)dh * sr.dt / (double)sr.n);
		++st;
		h += dh;
		if (h >= sr.n) {
			h = 0;
			++d;
			if constexpr (NF > 0) {
				if (d < sr.n_dump)
					s[8 + NF] = (rc.d[0] * s[8 + NF]) / rc.d[0];
			}
		}
	}
#pragma unroll
	for (int i = 0; i < 8 + NF; ++i)
		data[(size_t)i * n_rays + g] = s[i];
	if constexpr (NF > 0) {
#pragma unroll
		for (int i = 0; i < NF; ++i)
			data[(size_t)(8 + NF + i) * n_rays + g] = rc.d[i] * s[8 + NF];
	}
	V4<double> q = {s[0], s[1], s[2], s[3]};
	V4<double> u = {s[4], s[5], s[6], s[7]};
	info[g] = getuu(p.a, q, u);
	steps[g] += st;
}

/* SoA (device) <-> AoS (dump order) in the requested real size. */
template <class R>
__global__ __launch_bounds__(256) void k_soa_to_aos(const double *__restrict__ soa,
                                                    R *__restrict__ aos, size_t n_rays, int nd)
{
	size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= n_rays * (size_t)nd)
		return;
	size_t g = idx / nd, s = idx % nd;
	aos[idx] = (R)soa[s * n_rays + g];
}

template <class R>
__global__ __launch_bounds__(256) void k_aos_to_soa(const R *__restrict__ aos,
                                                    double *__restrict__ soa, size_t n_rays, int nd)
{
	size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (idx >= n_rays * (size_t)nd)
		return;
	size_t g = idx / nd, s = idx % nd;
	soa[s * n_rays + g] = (double)aos[idx];
}

template <class R>
__global__ __launch_bounds__(256) void k_cast(const double *__restrict__ in, R *__restrict__ out, size_t n)
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = (R)in[i];
}

__global__ __launch_bounds__(256) void k_info_f64(RayParams p, const double *__restrict__ data,
                                                  double *__restrict__ info, size_t n_rays)
{
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	V4<double> q = {data[g], data[n_rays + g], data[2 * n_rays + g], data[3 * n_rays + g]};
	V4<double> u = {data[4 * n_rays + g], data[5 * n_rays + g], data[6 * n_rays + g],
	                data[7 * n_rays + g]};
	info[g] = getuu(p.a, q, u);
}

/* Covariant view: out[g][8] = (q, k_mu) with k = down(q, u)
 * (sim-org/KS.cl:91-114), the oracle's og_down sequence. */
__global__ __launch_bounds__(256) void k_covariant(double a, const double *__restrict__ data,
                                                   double *__restrict__ out, size_t n_rays)
{
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	V4<double> q = {data[g], data[n_rays + g], data[2 * n_rays + g], data[3 * n_rays + g]};
	V4<double> u = {data[4 * n_rays + g], data[5 * n_rays + g], data[6 * n_rays + g],
	                data[7 * n_rays + g]};
	V4<double> k = down(a, q, u);
	double *o = out + (size_t)8 * g;
	o[0] = q.t; o[1] = q.x; o[2] = q.y; o[3] = q.z;
	o[4] = k.t; o[5] = k.x; o[6] = k.y; o[7] = k.z;
}

/* Boyer-Lindquist view of the final states (sim-org/KS.cl:63-149) in the
 * spheroidal convention of the metric, x + i y = (r + i a) e^{i phi}
 * sin(theta), and with u^r in the t and phi corrections (the derivative of
 * the position map of KS.cl:79-88).  Out: AoS [g][8] = (t, r, theta, phi,
 * u^t, u^r, u^theta, u^phi) in BL; t matched at r_match, phi at infinity.
 * Same expressions and order as the oracle; acos, atan2 and log are not
 * correctly rounded, so agreement is to ~1e-15 relative, not bitwise. */
__global__ __launch_bounds__(256) void k_to_bl(double a, double r_match,
                                               const double *__restrict__ data,
                                               double *__restrict__ out, size_t n_rays)
{
	size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= n_rays)
		return;
	const double q[4] = {data[g], data[n_rays + g], data[2 * n_rays + g], data[3 * n_rays + g]};
	const double u[4] = {data[4 * n_rays + g], data[5 * n_rays + g], data[6 * n_rays + g],
	                     data[7 * n_rays + g]};
	const double aa = a * a;
	const double RR = q[1] * q[1] + q[2] * q[2];
	const double zz = q[3] * q[3];
	/* position (og_getsphks, og_getbl) */
	const double kk0 = 0.5 * ((RR + zz) - aa);
	const double r = sqrt(sqrt(kk0 * kk0 + aa * zz) + kk0);
	const double h = sqrt(1.0 - aa), rp = 1.0 + h, rm = 1.0 - h;
	double *o = out + (size_t)8 * g;
	o[0] = q[0] + (rm * log((r - rm) / (r_match - rm)) - rp * log((r - rp) / (r_match - rp))) / h;
	o[1] = r;
	o[2] = acos(q[3] / r);
	o[3] = atan2(q[2] * r - q[1] * a, q[1] * r + q[2] * a) + ((0.5 * a) * log((r - rm) / (r - rp))) / h;
	/* velocity (og_getsphksu, og_getblu) */
	const double kk = 0.5 * ((RR + zz) - aa);
	const double dd = sqrt(kk * kk + aa * zz);
	const double rr = dd + kk;
	const double rs = sqrt(rr);
	const double d = (2.0 * rs) * dd;
	const double J1[4] = {0.0, (rr * q[1]) / d, (rr * q[2]) / d, ((rr + aa) * q[3]) / d};
	const double sz = sqrt(rr - zz);
	const double J2[4] = {((q[3] / rs) * J1[0]) / sz, ((q[3] / rs) * J1[1]) / sz,
	                      ((q[3] / rs) * J1[2]) / sz, ((q[3] / rs) * J1[3] - 1.0) / sz};
	const double c = a / (rr + aa);
	const double J3[4] = {c * J1[0], c * J1[1] + (-q[2]) / RR, c * J1[2] + q[1] / RR, c * J1[3]};
	const double ur = ((J1[0] * u[0] + J1[1] * u[1]) + J1[2] * u[2]) + J1[3] * u[3];
	const double ut = ((J2[0] * u[0] + J2[1] * u[1]) + J2[2] * u[2]) + J2[3] * u[3];
	const double up = ((J3[0] * u[0] + J3[1] * u[1]) + J3[2] * u[2]) + J3[3] * u[3];
	const double kk2 = 0.5 * (((q[1] * q[1] + q[2] * q[2]) + zz) - aa);
	const double rr2 = sqrt(kk2 * kk2 + aa * zz) + kk2;
	const double r2 = sqrt(rr2);
	const double D = (rr2 - 2.0 * r2) + aa;
	o[4] = u[0] - (ur * (2.0 * r2)) / D;
	o[5] = ur;
	o[6] = ut;
	o[7] = up - (ur * a) / D;
}

static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_to_bl(double a, double r_match, const double *data, double *out, size_t n_rays,
                        hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_to_bl, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st, a, r_match, data,
	                   out, n_rays);
	return hipGetLastError();
}

hipError_t launch_covariant(double a, const double *data, double *out, size_t n_rays,
                            hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_covariant, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st, a, data, out,
	                   n_rays);
	return hipGetLastError();
}

hipError_t launch_icond(const RayParams &p, double *data, double *info, int64_t *steps,
                        size_t n_rays, hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
#define L(NF) hipLaunchKernelGGL(k_icond<NF>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st, \
	                         p, data, info, steps, n_rays)
	GRAY_NF_SWITCH(p.n_freq, L)
#undef L
	return hipGetLastError();
}

hipError_t launch_evolve_f64(const RayParams &p, const StepRule &sr, double *data, double *info,
                             int64_t *steps, size_t n_rays, hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
#define L(NF) hipLaunchKernelGGL(k_evolve_f64<NF>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st, \
	                         p, sr, data, info, steps, n_rays)
	GRAY_NF_SWITCH(p.n_freq, L)
#undef L
	return hipGetLastError();
}

hipError_t launch_soa_to_aos(const double *soa, void *aos, int real_size, size_t n_rays, int nd,
                             hipStream_t st)
{
	size_t n = n_rays * (size_t)nd;
	if (n == 0)
		return hipSuccess;
	if (real_size == 8)
		hipLaunchKernelGGL(k_soa_to_aos<double>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   soa, (double *)aos, n_rays, nd);
	else
		hipLaunchKernelGGL(k_soa_to_aos<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   soa, (float *)aos, n_rays, nd);
	return hipGetLastError();
}

hipError_t launch_aos_to_soa(const void *aos, double *soa, int real_size, size_t n_rays, int nd,
                             hipStream_t st)
{
	size_t n = n_rays * (size_t)nd;
	if (n == 0)
		return hipSuccess;
	if (real_size == 8)
		hipLaunchKernelGGL(k_aos_to_soa<double>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   (const double *)aos, soa, n_rays, nd);
	else
		hipLaunchKernelGGL(k_aos_to_soa<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
		                   (const float *)aos, soa, n_rays, nd);
	return hipGetLastError();
}

hipError_t launch_cast(const double *in, void *out, int real_size, size_t n, hipStream_t st)
{
	if (n == 0)
		return hipSuccess;
	if (real_size == 8)
		return hipMemcpyAsync(out, in, n * sizeof(double), hipMemcpyDeviceToDevice, st);
	hipLaunchKernelGGL(k_cast<float>, dim3(blocks_for(n, 256)), dim3(256), 0, st,
	                   in, (float *)out, n);
	return hipGetLastError();
}

hipError_t launch_info_f64(const RayParams &p, const double *data, double *info, size_t n_rays,
                           hipStream_t st)
{
	if (n_rays == 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_info_f64, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st,
	                   p, data, info, n_rays);
	return hipGetLastError();
}

} // namespace gray
