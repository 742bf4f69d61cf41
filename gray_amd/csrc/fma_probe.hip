/*
 * fma_probe.hip -- gray-fma-probe: the attainable rate of non-packed wave64
 * FMAs on one GPU, FP32 (v_fmac_f32 / v_fma_f32) and FP64 (v_fma_f64), the
 * ceiling the ray-pool kernels are measured against when packed FP32 is off.
 *
 * Eight independent FMA chains per lane, 256-lane blocks (one wave per SIMD
 * each), 16 waves per SIMD, and again at the ray-pool kernels' own
 * occupancy (4 waves per SIMD for the FP32 pool, 3 for the serving pool, 2
 * for the plain FP64 kernel): at few waves a SIMD cannot hide an FMA's
 * latency even with eight independent chains.  One untimed launch, then one launch timed with HIP
 * events.  Prints one JSON line: TFLOP/s (FMA = 2 FLOP).
 *
 * FP32: every operand a vector register (the multiplier and addend differ
 * per lane); an FP32 instruction that reads a scalar register issues at
 * about half the rate, reported as fp32_sgpr_fma_tflops.  FP64: the
 * multiplier and addend uniform (scalar registers), the faster FP64 form
 * (three 64-bit vector operands issue slower); reported beside it as
 * fp64_vgpr_fma_tflops.
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class T, bool SGPR>
__global__ __launch_bounds__(256) void k_fma_chains(T *out, int iters, T a0, T b0)
{
	T x[8];
	/* per-lane multiplier and addend: vector registers */
	const T a = SGPR ? a0 : a0 + (T)(threadIdx.x * 1e-9);
	const T b = SGPR ? b0 : b0 + (T)(threadIdx.x * 1e-9);
#pragma unroll
	for (int k = 0; k < 8; ++k)
		x[k] = (T)(threadIdx.x * 1e-3) + (T)k;
	for (int i = 0; i < iters; ++i) {
#pragma unroll
		for (int k = 0; k < 8; ++k)
			x[k] = fma_t(x[k], a, b);
	}
	T s = 0;
#pragma unroll
	for (int k = 0; k < 8; ++k)
		s += x[k];
	out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

#define CK(e)                                                                      \
	do {                                                                       \
		hipError_t r_ = (e);                                               \
		if (r_ != hipSuccess) {                                            \
			fprintf(stderr, "gray-fma-probe: %s: %s\n", #e, hipGetErrorString(r_)); \
			exit(1);                                                   \
		}                                                                  \
	} while (0)

template <class T, bool SGPR>
double rate(int blocks, int iters)
{
	T *out;
	CK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(T)));
	hipEvent_t e0, e1;
	CK(hipEventCreate(&e0));
	CK(hipEventCreate(&e1));
	hipLaunchKernelGGL((k_fma_chains<T, SGPR>), dim3(blocks), dim3(256), 0, 0, out, 1000, (T)0.999, (T)1e-3);
	CK(hipGetLastError());
	CK(hipDeviceSynchronize());
	CK(hipEventRecord(e0, 0));
	hipLaunchKernelGGL((k_fma_chains<T, SGPR>), dim3(blocks), dim3(256), 0, 0, out, iters, (T)0.999, (T)1e-3);
	CK(hipGetLastError());
	CK(hipEventRecord(e1, 0));
	CK(hipEventSynchronize(e1));
	float ms = 0.f;
	CK(hipEventElapsedTime(&ms, e0, e1));
	CK(hipEventDestroy(e0));
	CK(hipEventDestroy(e1));
	CK(hipFree(out));
	return 2.0 * 8.0 * (double)iters * (double)blocks * 256.0 / ((double)ms * 1e-3) / 1e12;
}

} // namespace

int main()
{
	int cus = 0;
	CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
	/* a block of 4 waves puts one wave on each SIMD of a CU */
	const double f32 = rate<float, false>(16 * cus, 200000);
	const double f32s = rate<float, true>(16 * cus, 200000);
	const double f32_3 = rate<float, false>(3 * cus, 200000);
	const double f32_4 = rate<float, false>(4 * cus, 200000);
	const double f64 = rate<double, true>(16 * cus, 100000);
	const double f64v = rate<double, false>(16 * cus, 100000);
	const double f64_2 = rate<double, true>(2 * cus, 100000);
	printf("{\"fp32_fma_tflops\": %.3f, \"fp32_sgpr_fma_tflops\": %.3f, "
	       "\"fp32_fma_tflops_3_waves\": %.3f, \"fp32_fma_tflops_4_waves\": %.3f, "
	       "\"fp64_fma_tflops\": %.3f, "
	       "\"fp64_vgpr_fma_tflops\": %.3f, \"fp64_fma_tflops_2_waves\": %.3f, \"cus\": %d, "
	       "\"waves_per_simd\": 16, \"chains_per_lane\": 8}\n",
	       f32, f32s, f32_3, f32_4, f64, f64v, f64_2, cus);
	return 0;
}
